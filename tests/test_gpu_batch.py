"""Batched frames (fsr_batch_run): the eager run and the graph replays give the
same bits as the per-frame entry point, frame by frame; the autotuner keeps
results unchanged (launch shapes do not change any value)."""
import ctypes

import numpy as np
import pytest

from fluidsurfacereconstructiongpu_amd import _lib
from fluidsurfacereconstructiongpu_amd import datasets as ds
from fluidsurfacereconstructiongpu_amd.reconstructor import Batch, Reconstructor, stats_dict

pytestmark = pytest.mark.gpu


def small_frames(method, count=6):
    kw = dict(n_min=3e3, n_max=4e4, res_min=40, res_max=96)
    if method == "yu13":
        kw.update(method="yu13", iso=0.0, scale=1.3)
    return [ds.random_shapes_frame(i, **kw) for i in range(count)]


def per_frame(method, frames, aniso=None):
    rec = Reconstructor(method, aniso=aniso)
    out = []
    for fr in frames:
        res = rec.run(fr, tri_capacity=max(3 * 1024, 4 * fr.n))
        T = res.stats["num_mesh_vertices"]
        out.append((res.phi.cpu().numpy().copy(), res.tri_pos[:T].cpu().numpy().copy(),
                    res.tri_nrm[:T].cpu().numpy().copy(), res.stats))
    return out


def check_same(batch, ref):
    for i, (b, (phi, tp, tn, st)) in enumerate(zip(batch.bufs, ref)):
        sb = stats_dict(b["stats"])
        assert sb == st, i
        T = st["num_mesh_vertices"]
        assert np.array_equal(b["phi"].cpu().numpy(), phi), i
        assert np.array_equal(b["tri"][:T].cpu().numpy(), tp), i
        assert np.array_equal(b["nrm"][:T].cpu().numpy(), tn), i


@pytest.mark.parametrize("method", ["zb05", "yu13"])
def test_batch_equals_per_frame(torch_cuda, method):
    torch = torch_cuda
    print(f"\n[gpu case] batch {method}", flush=True)
    frames = small_frames(method)
    aniso = ds.yu13_params(frames[0], True) if method == "yu13" else None
    ref = per_frame(method, frames, aniso)
    batch = Batch(method, frames, lanes=3, aniso=aniso, cap_factor=2.0)
    batch.run()                      # eager + capture
    torch.cuda.synchronize()
    if batch.grow_overflowed():      # a frame's soup did not fit: new descriptors, recapture
        batch.run()
        torch.cuda.synchronize()
    assert not batch.grow_overflowed()
    check_same(batch, ref)
    for b in batch.bufs:             # poison the outputs, replay the graph twice
        b["phi"].fill_(123.0)
        b["tri"].fill_(-5.0)
    for _ in range(2):
        batch.run()
        torch.cuda.synchronize()
        check_same(batch, ref)
        ms = batch.frame_ms()
        assert np.all(ms > 0) and np.all(ms < 1e4)
    batch.close()


def test_autotune_keeps_results(torch_cuda):
    torch = torch_cuda
    print("\n[gpu case] autotune", flush=True)
    fr = ds.jittered_lattice(n_side=24, grid_res=80)
    rec = Reconstructor("zb05")
    res = rec.run(fr)
    phi0 = res.phi.cpu().numpy().copy()
    T = res.stats["num_mesh_vertices"]
    tri0 = res.tri_pos[:T].cpu().numpy().copy()
    batch = Batch("zb05", [fr], lanes=1, ctx=rec.ctx)
    choice = (ctypes.c_uint32 * 4)()
    ap = rec.aniso_param()
    _lib.check(rec.ctx.lib.fsr_autotune(rec.ctx.handle, 0, ctypes.byref(ap), batch.descs, 3,
                                        choice, ctypes.c_void_p(0)))
    # 0 = the kernels' resident count; otherwise a limit below it
    assert all(c in (0, 2, 3, 4) for c in list(choice)[:2] + [choice[3]])
    res2 = rec.run(fr)
    assert np.array_equal(res2.phi.cpu().numpy(), phi0)
    assert np.array_equal(res2.tri_pos[:T].cpu().numpy(), tri0)
    batch.close()


def test_batch_autotune_keeps_results(torch_cuda):
    """fsr_batch_autotune changes only the lanes' launch shapes: the batch after
    it gives the per-frame bits."""
    torch = torch_cuda
    print("\n[gpu case] batch autotune", flush=True)
    frames = small_frames("zb05", count=4)
    ref = per_frame("zb05", frames)
    batch = Batch("zb05", frames, lanes=2, cap_factor=2.0)
    choice = ctypes.c_uint32(99)
    ap = batch.rec.aniso_param()
    _lib.check(batch.ctx.lib.fsr_batch_autotune(batch.handle, 0, ctypes.byref(ap), batch.descs,
                                                len(frames), 2, ctypes.byref(choice),
                                                ctypes.c_void_p(0)))
    torch.cuda.synchronize()
    assert choice.value in (0, 1, 2, 3, 4)
    if batch.grow_overflowed():
        batch.run()
        torch.cuda.synchronize()
    batch.run()
    torch.cuda.synchronize()
    check_same(batch, ref)
    batch.close()

