"""GPU adjoint (fsr_mc_vertex_backward, fsr_field_zb05_backward and the
torch.autograd binding) against the FP64 oracle.

Tolerances:
  * vertex term, per band vertex: |g_gpu - g_ref| <= 1e-4 |g_ref| + 1e-6 max |g_ref|
    (FP32 evaluation of (iso - phi) / D^2 against FP64 on the same field);
  * particle gradient: per particle |g_gpu - g_ref| <= 1e-4 max |g_ref| and
    ||g_gpu - g_ref|| / ||g_ref|| <= 1e-4 (FP32 sums over ~10^3 vertex terms of
    either sign, from a random upstream, against FP64);
  * against FP64 central differences of the oracle's loss (eps 1e-6, 64 seeded
    particles with a non-negligible gradient): relative error < 1e-3.
The oracle side is fed the GPU's field and soup, so both sides differentiate
the same topology.
"""
import ctypes

import numpy as np
import pytest

import oracle
from fluidsurfacereconstructiongpu_amd import datasets as ds
from fluidsurfacereconstructiongpu_amd import _lib
from tests import parity

pytestmark = pytest.mark.gpu

VERTEX_RTOL = 1e-4
VERTEX_FLOOR = 1e-6
PART_MAX = 1e-4
PART_NORM = 1e-4
FD_RTOL = 1e-3


def _frames():
    return {
        "ball4k": lambda: ds.lattice_ball(n_target=4000, grid_res=48),
        "ball32k": lambda: ds.lattice_ball(n_target=32_000, grid_res=96),
        "lattice40": lambda: ds.jittered_lattice(),
    }


CASES = [("ball4k", "zb05"), ("ball4k", "akinci12"), ("ball32k", "zb05"),
         ("lattice40", "zb05"), ("lattice40", "akinci12")]


def oracle_chain(ref, fr, method, phi_gpu, g_phi, g_soup):
    """FP64: vertex term then particle gradient (caller order)."""
    g_tot = oracle.mc_vertex_backward(ref.index, ref.band, phi_gpu.astype(np.float64), fr.sc,
                                      np.float32(fr.iso), g_soup.astype(np.float64),
                                      cpu=(method == "akinci12"),
                                      g_phi=g_phi.astype(np.float64))
    g_band = np.zeros_like(g_tot)
    g_band[ref.index] = g_tot[ref.index]
    h = np.float32(fr.h)
    gs = oracle.field_zb05_backward(ref.index, ref.hs["pos"], ref.hs["ranges"], fr.sp, fr.sc,
                                    h * h, g_band)
    g = np.empty_like(gs)
    g[ref.hs["perm"]] = gs
    return g_tot, g


def particle_errors(g_gpu, g_ref):
    d = np.linalg.norm(g_gpu.astype(np.float64) - g_ref, axis=1)
    scale = np.linalg.norm(g_ref, axis=1).max()
    return float(d.max() / scale), float(np.linalg.norm(d) / np.linalg.norm(g_ref))


@pytest.fixture(scope="module", params=CASES, ids=["%s-%s" % c for c in CASES])
def case(request, torch_cuda):
    from fluidsurfacereconstructiongpu_amd.autograd import SurfaceField

    name, method = request.param
    print(f"\n[gpu adjoint case] {name} {method}", flush=True)
    torch = torch_cuda
    fr = _frames()[name]()
    sf = SurfaceField(fr, method)
    pos = torch.from_numpy(fr.pos).cuda()
    phi, soup = sf.forward_frame(pos)
    ref = parity.Stage(fr, method)
    return dict(fr=fr, method=method, sf=sf, pos=pos, phi=phi, soup=soup, ref=ref,
                torch=torch)


def _backward_parts(case, g_phi, g_soup):
    torch, sf = case["torch"], case["sf"]
    lib, h = sf.ctx.lib, sf.ctx.handle
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    gp = torch.from_numpy(g_phi).cuda().contiguous()
    gs = torch.from_numpy(g_soup).cuda().contiguous()
    _lib.check(lib.fsr_mc_vertex_backward(h, ctypes.c_void_p(gs.data_ptr()),
                                          ctypes.c_void_p(gp.data_ptr()), s))
    out = torch.empty((case["fr"].n, 3), dtype=torch.float32, device="cuda")
    _lib.check(lib.fsr_field_zb05_backward(h, ctypes.c_void_p(gp.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), s))
    torch.cuda.synchronize()
    return gp.cpu().numpy(), out.cpu().numpy()


def test_band_matches_oracle(case):
    st = case["sf"].stats
    assert st["num_band_vertices"] == case["ref"].index.shape[0]
    assert st["num_mesh_vertices"] == case["soup"].shape[0] > 0


def test_backward_random_upstream(case):
    fr, ref, method = case["fr"], case["ref"], case["method"]
    rng = np.random.default_rng(11)
    K = case["soup"].shape[0]
    g_phi = rng.standard_normal(fr.sc.count).astype(np.float32)
    g_soup = rng.standard_normal((K, 3)).astype(np.float32)
    gp, gpos = _backward_parts(case, g_phi, g_soup)
    g_tot, g_ref = oracle_chain(ref, fr, method, case["phi"].cpu().numpy(), g_phi, g_soup)
    # vertex term at band vertices; other vertices untouched
    b = ref.index
    diff = np.abs(gp[b].astype(np.float64) - g_tot[b])
    lim = VERTEX_RTOL * np.abs(g_tot[b]) + VERTEX_FLOOR * np.abs(g_tot[b]).max()
    assert int((diff > lim).sum()) == 0, float((diff / np.maximum(lim, 1e-300)).max())
    mask = np.ones(fr.sc.count, bool)
    mask[b] = False
    assert np.array_equal(gp[mask], g_phi[mask])
    # particle gradient from the GPU's own g_phi
    _, g_ref2 = oracle_chain(ref, fr, method, case["phi"].cpu().numpy(), gp,
                             np.zeros_like(g_soup))
    worst, rel = particle_errors(gpos, g_ref2)
    print(f"  particle gradient: worst {worst:.2e} (of max), norm {rel:.2e}")
    assert worst <= PART_MAX and rel <= PART_NORM, (worst, rel)


def test_backward_deterministic(case):
    rng = np.random.default_rng(3)
    K = case["soup"].shape[0]
    g_phi = rng.standard_normal(case["fr"].sc.count).astype(np.float32)
    g_soup = rng.standard_normal((K, 3)).astype(np.float32)
    a = _backward_parts(case, g_phi, g_soup)
    b = _backward_parts(case, g_phi, g_soup)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_autograd_sphere_loss(case):
    from fluidsurfacereconstructiongpu_amd.autograd import sphere_loss

    torch, fr, sf = case["torch"], case["fr"], case["sf"]
    x = case["pos"].clone().requires_grad_(True)
    phi, soup = sf(x)
    loss = sphere_loss(phi, soup, fr)
    loss.backward()
    g_gpu = x.grad.cpu().numpy()
    # the same upstream in FP64 from the GPU's outputs
    phi_n = phi.detach().cpu().numpy()
    V, K = fr.sc.count, soup.shape[0]
    ph = phi.detach().clone().requires_grad_(True)
    so = soup.detach().clone().requires_grad_(True)
    sphere_loss(ph, so, fr).backward()
    g_phi_up = ph.grad.cpu().numpy()
    g_soup_up = so.grad.cpu().numpy()
    _, g_ref = oracle_chain(case["ref"], fr, case["method"], phi_n, g_phi_up, g_soup_up)
    worst, rel = particle_errors(g_gpu, g_ref)
    print(f"  autograd vs oracle chain: worst {worst:.2e}, norm {rel:.2e}")
    assert worst <= PART_MAX and rel <= PART_NORM
    assert V > 0 and K > 0


def test_stale_graph_raises(case):
    torch, sf = case["torch"], case["sf"]
    x = case["pos"].clone().requires_grad_(True)
    phi, soup = sf(x)
    loss = phi.sum() + soup.sum()
    sf.forward_frame(case["pos"])  # a newer frame replaces the adjoint state
    with pytest.raises(RuntimeError):
        loss.backward()


@pytest.mark.parametrize("method", ["yu13"])
def test_no_adjoint_for_yu13(torch_cuda, method):
    from fluidsurfacereconstructiongpu_amd.reconstructor import Context, Reconstructor

    fr = ds.jittered_lattice(n_side=12, grid_res=48)
    ctx = Context(0)
    _lib.check(ctx.lib.fsr_set_adjoint(ctx.handle, 1))
    rec = Reconstructor(method, ctx=ctx)
    rec.run(fr)
    g = torch_cuda.zeros(fr.sc.count, dtype=torch_cuda.float32, device="cuda")
    out = torch_cuda.zeros((fr.n, 3), dtype=torch_cuda.float32, device="cuda")
    st = ctx.lib.fsr_field_zb05_backward(ctx.handle, ctypes.c_void_p(g.data_ptr()),
                                         ctypes.c_void_p(out.data_ptr()), None)
    assert st == 6  # FSR_ERR_UNSUPPORTED
    ctx.close()


def test_c5_ball256k_against_fd(torch_cuda):
    """The adjoint workload at full size: 256,000-particle jittered lattice ball,
    192^3 grid, sphere loss; the GPU gradient against the FP64 oracle chain on
    the GPU's topology, and against FP64 central differences of the oracle's
    loss on 64 seeded particles."""
    from fluidsurfacereconstructiongpu_amd.autograd import SurfaceField, sphere_loss

    torch = torch_cuda
    print("\n[gpu adjoint case] ball256k 192^3 zb05", flush=True)
    fr = ds.lattice_ball()
    sf = SurfaceField(fr, "zb05")
    x = torch.from_numpy(fr.pos).cuda().requires_grad_(True)
    phi, soup = sf(x)
    loss = sphere_loss(phi, soup, fr)
    loss.backward()
    g_gpu = x.grad.cpu().numpy().astype(np.float64)
    print("  gpu done, loss %.6e" % float(loss.detach()), flush=True)
    A = oracle.AdjointWorkload("zb05", fr)
    print("  oracle workload B=%d K=%d (gpu B=%d K=%d) loss %.6e" % (
        A.B, A.K, sf.stats["num_band_vertices"], sf.stats["num_mesh_vertices"], A.loss()),
        flush=True)
    assert A.B == sf.stats["num_band_vertices"]
    g_orc = A.grad()
    mag = np.linalg.norm(g_orc, axis=1)
    cand = np.nonzero(mag > 1e-3 * mag.max())[0]
    sel = np.random.default_rng(2024).choice(cand, 64, replace=False)
    fd = np.array([[A.fd(j, a, 1e-6) for a in range(3)] for j in sel])
    per_orc = np.linalg.norm(fd - g_orc[sel], axis=1) / np.linalg.norm(fd, axis=1)
    per_gpu = np.linalg.norm(fd - g_gpu[sel], axis=1) / np.linalg.norm(fd, axis=1)
    print("  FD rel err: oracle max %.2e, gpu max %.2e" % (per_orc.max(), per_gpu.max()),
          flush=True)
    assert per_orc.max() < FD_RTOL
    assert per_gpu.max() < FD_RTOL
    A.close()
