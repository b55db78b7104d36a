"""The stage entry points (one per reference launcher) driven in the reference
reconstructors' own call order (GPU/ReconstructorGPUOursZB05.cpp:22-78 with
GPU/ReconstructorGPUOurs.cpp:164-265; GPU/ReconstructorGPUAkinci12.cpp:34-84;
GPU/ReconstructorGPUOursYu13.cpp:24-64), with the reference's buffer formats
(dense u32 flag grids, scans with host totals, float3 arrays).  Their results
must equal the whole-frame entry point bit for bit (the frame entry point is
itself checked against the oracle in test_gpu_zb05.py / test_gpu_yu13.py)."""
import ctypes

import numpy as np
import pytest

from fluidsurfacereconstructiongpu_amd import _lib
from fluidsurfacereconstructiongpu_amd import datasets as ds
from fluidsurfacereconstructiongpu_amd.reconstructor import Reconstructor, grid_info

pytestmark = pytest.mark.gpu


class Stages:
    def __init__(self, torch, rec):
        self.t = torch
        self.lib = rec.ctx.lib
        self.ctx = rec.ctx.handle.value
        self.dev = rec.device
        self.keep = []

    def p(self, x):
        if hasattr(x, "data_ptr"):
            return x.data_ptr()
        self.keep.append(x)
        return ctypes.addressof(x)

    def call(self, name, *args):
        conv = [self.p(a) if not isinstance(a, (int, type(None), ctypes.c_uint32)) and
                not hasattr(a, "_type_") else a for a in args]
        st = getattr(self.lib, name)(self.ctx, *conv, None)
        _lib.check(st)

    def u32(self, n):
        return self.t.zeros(max(int(n), 1), dtype=self.t.int32, device=self.dev)

    def f32(self, n, fill=0.0):
        return self.t.full((max(int(n), 1),), fill, dtype=self.t.float32, device=self.dev)

    def scan(self, out, inp, n):
        total = ctypes.c_uint32()
        st = self.lib.fsr_thrust_exclusive_prefix_sum_scan(self.ctx, out.data_ptr(),
                                                            inp.data_ptr(), int(n),
                                                            ctypes.byref(total), None)
        _lib.check(st)
        return total.value


def run_frame(fr, method, aniso=None):
    rec = Reconstructor(method, aniso=aniso)
    res = rec.run(fr)
    T = res.stats["num_mesh_vertices"]
    return rec, res.stats, res.phi.cpu().numpy().copy(), res.tri_pos[:T].cpu().numpy().copy(), \
        res.tri_nrm[:T].cpu().numpy().copy()


def mc_stages(S, sim, scg, sv, B, phi, is_surf, akinci=False):
    valid = S.u32(B)
    if akinci:
        S.call("fsr_detection_of_valid_surface_cubes_for_akinci", sv, B, phi, scg, valid, is_surf,
               sim)
        return valid, None, None
    nverts = S.u32(B)
    S.call("fsr_detect_valid_surface_cubes", sv, B, phi, scg, valid, nverts, is_surf, sim)
    valid_scan, nverts_scan = S.u32(B), S.u32(B)
    Q = S.scan(valid_scan, valid, B)
    T3 = S.scan(nverts_scan, nverts, B)
    valid_idx = S.u32(Q)
    S.call("fsr_compact_valid_surface_cubes", valid_idx, valid_scan, valid, B)
    pos, nrm = S.f32(3 * T3), S.f32(3 * T3)
    S.call("fsr_generate_triangles", sv, valid_idx, Q, scg, nverts_scan, phi, pos, nrm, T3, sim)
    return valid, pos[:3 * T3].view(-1, 3), nrm[:3 * T3].view(-1, 3)


@pytest.mark.parametrize("method", ["zb05", "akinci12"])
def test_reference_stage_sequence_zb05_akinci(torch_cuda, method):
    torch = torch_cuda
    print(f"\n[gpu case] stage sequence {method}", flush=True)
    fr = ds.jittered_lattice(n_side=24, grid_res=80)
    rec, st, phi_f, pos_f, nrm_f = run_frame(fr, method)
    S = Stages(torch, rec)
    sim = rec.params(fr)
    spg, scg = grid_info(fr.sp), grid_info(fr.sc)
    n, C, V = fr.n, fr.sp.count, fr.sc.count
    pos = torch.from_numpy(fr.pos).cuda().contiguous()
    dens = torch.from_numpy(fr.dens).cuda()
    ranges, flag = S.u32(2 * C), S.f32(C)
    S.call("fsr_spatial_grid_building", pos, dens, n, ranges, flag, spg)
    flags = S.u32(n)
    if method == "zb05":
        S.call("fsr_extraction_of_surface_particles", sim, flag, flags, n, ranges, spg)
    else:
        S.call("fsr_extraction_of_surface_particles_for_akinci", sim, pos, dens, n, ranges, spg,
               flags)
    is_surf = S.u32(V)
    S.call("fsr_estimation_of_surface_vertices", sim, scg, pos, flags, n, ranges, is_surf)
    is_scan = S.u32(V)
    B = S.scan(is_scan, is_surf, V)
    assert B == st["num_band_vertices"]
    sv = S.u32(B)
    S.call("fsr_compact_surface_vertex", sv, is_scan, is_surf, V)
    phi = S.f32(V, 0.0)
    S.call("fsr_update_scalar_grid_values_compacted", sv, B, ranges, pos, n, phi, spg, scg, sim)
    torch.cuda.synchronize()
    assert np.array_equal(pos.cpu().numpy(), rec.pos_sorted.cpu().numpy())
    assert np.array_equal(flags.cpu().numpy().view(np.uint32), rec.buffer("surface_flags", np.uint32, n))
    assert np.array_equal(sv.cpu().numpy().view(np.uint32), rec.buffer("band_index", np.uint32, B))
    assert np.array_equal(phi[:V].cpu().numpy(), phi_f)
    valid, tp, tn = mc_stages(S, sim, scg, sv, B, phi, is_surf, akinci=method == "akinci12")
    torch.cuda.synchronize()
    nv_f = rec.buffer("cube_nverts", np.uint32, B)
    assert np.array_equal(valid[:B].cpu().numpy().view(np.uint32), (nv_f > 0).astype(np.uint32))
    if method == "zb05":
        assert np.array_equal(tp.cpu().numpy(), pos_f)
        assert np.array_equal(tn.cpu().numpy(), nrm_f)


@pytest.mark.parametrize("config_mode", [True, False])
def test_reference_stage_sequence_yu13(torch_cuda, config_mode):
    torch = torch_cuda
    print(f"\n[gpu case] stage sequence yu13 config_mode={config_mode}", flush=True)
    fr = ds.jittered_lattice(n_side=20, grid_res=80, scale=1.3, iso=0.0, method="yu13",
                             cell=0.004)
    aniso = ds.yu13_params(fr, config_mode)
    rec, st, phi_f, pos_f, nrm_f = run_frame(fr, "yu13", aniso)
    S = Stages(torch, rec)
    sim = rec.params(fr)
    ap = rec.aniso_param()
    spg, scg = grid_info(fr.sp), grid_info(fr.sc)
    n, C, V = fr.n, fr.sp.count, fr.sc.count
    pos = torch.from_numpy(fr.pos).cuda().contiguous()
    dens = torch.from_numpy(fr.dens).cuda()
    ranges, flag = S.u32(2 * C), S.f32(C)
    S.call("fsr_spatial_grid_building", pos, dens, n, ranges, flag, spg)
    flags, involve = S.u32(n), S.u32(C)
    S.call("fsr_extraction_of_surface_and_involve_particles", sim, flag, involve, flags, n,
           ranges, spg)
    is_surf = S.u32(V)
    S.call("fsr_estimation_of_surface_vertices", sim, scg, pos, flags, n, ranges, is_surf)
    mean, smooth = S.f32(3 * n), S.f32(3 * n)
    S.call("fsr_calculation_of_mean_and_smooth_particles", sim, pos, n, mean, smooth, ranges, spg)
    inv_scan, is_scan = S.u32(C), S.u32(V)
    I = S.scan(inv_scan, involve, C)
    B = S.scan(is_scan, is_surf, V)
    assert I == st["num_involved_particles"] and B == st["num_band_vertices"]
    sv, inv_idx = S.u32(B), S.u32(I)
    S.call("fsr_compactation_of_surface_vertices_and_particles", is_surf, is_scan, V, involve,
           inv_scan, ranges, C, sv, inv_idx)
    G = S.f32(9 * I)
    S.call("fsr_calculation_of_transform_matrices_for_particles", sim, ap, spg, mean, pos, n,
           ranges, inv_idx, I, G)
    phi = torch.full((V,), 0.0, dtype=torch.float32, device=rec.device)
    phi.view(torch.int32).fill_(0x64646464)
    S.call("fsr_computation_of_scalar_field_grid", sim, B, scg, spg, G, smooth, dens, n, ranges,
           sv, involve, inv_scan, phi)
    torch.cuda.synchronize()
    assert np.array_equal(inv_idx[:I].cpu().numpy().view(np.uint32),
                          rec.buffer("involved_index", np.uint32, I))
    assert np.array_equal(G[:9 * I].cpu().numpy(), rec.buffer("g_matrices", np.float32, 9 * I))
    sm_f = rec.buffer("smooth_pos", np.float32, 4 * n).reshape(n, 4)[:, :3]
    assert np.array_equal(smooth[:3 * n].cpu().numpy().reshape(n, 3), sm_f)
    # The field: the frame culls each candidate with its own support radius
    # before it enters a 64-candidate batch; the launcher, given only G, does
    # not.  Every vertex sums its batches as partial sums, so the two group the
    # same (exact-zero) terms differently and round apart by a few ulp.
    phi_s = phi.cpu().numpy()
    fill = np.uint32(0x64646464)
    band = phi_f.view(np.uint32) != fill
    assert np.array_equal(phi_s.view(np.uint32)[~band], phi_f.view(np.uint32)[~band])
    d = np.abs(phi_s[band].astype(np.float64) - phi_f[band])
    assert np.all(d <= 1e-6 * np.maximum(np.abs(phi_f[band]), 1.0)), float(d.max())
    # the marching cubes stages on the frame's field give the frame's mesh bits
    phi.copy_(torch.from_numpy(phi_f).to(rec.device))
    valid, tp, tn = mc_stages(S, sim, scg, sv, B, phi, is_surf)
    torch.cuda.synchronize()
    assert np.array_equal(tp.cpu().numpy(), pos_f)
    assert np.array_equal(tn.cpu().numpy(), nrm_f)


def tri_table_from_packed():
    """int[256][16] triangle table (the reference's triTable layout, -1 after
    the last edge) unpacked from the oracle's packed table, which
    tests/test_mc_tables.py checks entry by entry against the reference's."""
    import oracle

    out = np.full((256, 16), -1, np.int32)
    for c, v in enumerate(oracle.mc_tri_packed()):
        v = int(v)
        for k in range(16):
            e = (v >> (4 * k)) & 0xF
            if e == 0xF:
                break
            out[c, k] = e
    return out


def test_bind_textures_checks_table(torch_cuda):
    torch = torch_cuda
    rec = Reconstructor("zb05")
    lib, ctx = rec.ctx.lib, rec.ctx.handle.value
    tri = torch.from_numpy(tri_table_from_packed().ravel()).cuda()
    assert lib.fsr_bind_textures(ctx, None, tri.data_ptr(), None, None) == 0
    bad = tri.clone()
    bad[16 * 5 + 1] = 7
    assert lib.fsr_bind_textures(ctx, None, bad.data_ptr(), None, None) == 1
    assert lib.fsr_bind_textures(ctx, None, None, None, None) == 0
