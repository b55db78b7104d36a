"""GPU entry points of the two launchers the reference defines but never calls,
against the oracle: estimateSurfaceVertices bit-exact; the isotropic TC01 field
within |dphi| <= 1e-5 max(|phi|, 1) (FP32 sums of up to ~2000 terms in [0, 1/4])."""
import ctypes

import numpy as np
import pytest

import oracle
from fluidsurfacereconstructiongpu_amd import _lib
from fluidsurfacereconstructiongpu_amd import datasets as ds
from fluidsurfacereconstructiongpu_amd.reconstructor import Context, grid_info, sim_param
from tests import parity

pytestmark = pytest.mark.gpu


def P(t):
    return ctypes.c_void_p(t.data_ptr())


def test_estimate_surface_vertices(torch_cuda):
    torch = torch_cuda
    print("\n[gpu case] estimate_surface_vertices", flush=True)
    rng = np.random.default_rng(9)
    dg = ds.Grid(np.zeros(3, np.float32), np.full(3, 13, np.float32), np.float32(1.0),
                 (13, 11, 9))
    ctx = Context(0)
    for ratio, ext in ((2, 4), (3, 3), (1, 0)):
        sg_res = tuple(int(r * ratio + 1) for r in dg.res)
        sg = ds.Grid(np.zeros(3, np.float32), np.full(3, 13, np.float32), np.float32(1.0 / ratio),
                     sg_res)
        dens = (rng.uniform(size=dg.count) < 0.3).astype(np.float32)
        ref = oracle.estimate_surface_vertices(dens, dg, ratio, ext, sg)
        d = torch.from_numpy(dens).cuda()
        out = torch.zeros(sg.count, dtype=torch.int32, device="cuda")
        sim = sim_param(dict(sc_sp_grid_res_ratio=ratio))
        gdg, gsg = grid_info(dg), grid_info(sg)
        _lib.check(ctx.lib.fsr_estimate_surface_vertices(ctx.handle, ctypes.byref(sim), P(d),
                                                         ctypes.byref(gdg), ext, P(out),
                                                         ctypes.byref(gsg), None))
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().view(np.uint32), ref)
    ctx.close()


@pytest.mark.parametrize("n_side,res", [(16, 64), (24, 96)])
def test_tc01_field(torch_cuda, n_side, res):
    torch = torch_cuda
    print(f"\n[gpu case] tc01 lattice{n_side} {res}", flush=True)
    fr = ds.jittered_lattice(n_side=n_side, grid_res=res)
    st = parity.Stage(fr, "zb05")
    h = np.float32(fr.h)
    ref = oracle.field_tc01(st.index, st.hs["pos"], st.hs["ranges"], fr.sp, fr.sc, h * h)
    ctx = Context(0)
    pos = torch.from_numpy(np.ascontiguousarray(st.hs["pos"])).cuda()
    ranges = torch.from_numpy(st.hs["ranges"].view(np.int32).ravel().copy()).cuda()
    idx = torch.from_numpy(st.index.view(np.int32)).cuda()
    phi = torch.zeros(fr.sc.count, dtype=torch.float32, device="cuda")
    sim = sim_param(ds.sim_param_fields(fr))
    gsp, gsc = grid_info(fr.sp), grid_info(fr.sc)
    _lib.check(ctx.lib.fsr_update_scalar_grid_values_tc01(
        ctx.handle, P(idx), st.index.shape[0], P(ranges), P(pos), fr.n, P(phi), ctypes.byref(gsp),
        ctypes.byref(gsc), ctypes.byref(sim), None))
    torch.cuda.synchronize()
    got = phi.cpu().numpy().astype(np.float64)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    assert err.max() <= 1e-5, err.max()
    assert np.count_nonzero(ref[st.index]) > 0
    ctx.close()
