"""fsr_exclusive_scan_noalloc (the allocation-free scan, scratch in a fixed
per-device array) against fsr_thrust_exclusive_prefix_sum_scan: bit-exact
outputs and totals on two calls in a row, the second on a smaller input, so the
second call runs on scratch that still holds the first call's partials."""
import ctypes

import numpy as np
import pytest

from fluidsurfacereconstructiongpu_amd import _lib
from fluidsurfacereconstructiongpu_amd.reconstructor import Context

pytestmark = pytest.mark.gpu


def _scan_pair(torch, ctx, host):
    lib, h = ctx.lib, ctx.handle
    n = host.shape[0]
    x = torch.from_numpy(host.view(np.int32)).cuda()
    out_ref = torch.full((n + 4,), -1, dtype=torch.int32, device="cuda")
    out_na = torch.full((n + 4,), -1, dtype=torch.int32, device="cuda")
    tot = torch.full((2,), -1, dtype=torch.int32, device="cuda")
    host_total = ctypes.c_uint32(0)
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.fsr_thrust_exclusive_prefix_sum_scan(h, out_ref.data_ptr(), x.data_ptr(), n,
                                                        ctypes.byref(host_total), stream))
    _lib.check(lib.fsr_exclusive_scan_noalloc(h, out_na.data_ptr(), x.data_ptr(), n,
                                              tot.data_ptr(), stream))
    torch.cuda.synchronize()
    ref = out_ref.cpu().numpy().view(np.uint32)
    got = out_na.cpu().numpy().view(np.uint32)
    expect = np.zeros(n, np.uint64)
    expect[1:] = np.cumsum(host[:-1], dtype=np.uint64)
    assert np.array_equal(ref[:n], (expect & 0xFFFFFFFF).astype(np.uint32))
    assert np.array_equal(got[:n], ref[:n])
    # nothing written past the end
    assert np.all(got[n:] == 0xFFFFFFFF)
    t = tot.cpu().numpy().view(np.uint32)
    assert t[0] == host_total.value == np.uint32(np.sum(host, dtype=np.uint64) & 0xFFFFFFFF)
    assert t[1] == 0xFFFFFFFF


def test_noalloc_scan_matches_arena_scan_twice(torch_cuda):
    torch = torch_cuda
    ctx = Context(0)
    try:
        rng = np.random.default_rng(20261203)
        # first call: 1024 workgroups, 5 tiles each; second: fewer workgroups
        big = rng.integers(0, 1 << 12, size=5 * 1024 * 4096 + 77, dtype=np.uint32)
        small = rng.integers(0, 1 << 20, size=300_001, dtype=np.uint32)
        print(f"\n[gpu case] noalloc scan n={big.shape[0]} then n={small.shape[0]}", flush=True)
        _scan_pair(torch, ctx, big)
        _scan_pair(torch, ctx, small)
        _scan_pair(torch, ctx, np.ones(1, np.uint32))
    finally:
        ctx.close()


def test_noalloc_scan_rejects_oversize_and_null(torch_cuda):
    ctx = Context(0)
    try:
        lib, h = ctx.lib, ctx.handle
        assert lib.fsr_exclusive_scan_noalloc(h, None, None, (1 << 30) + 1, None, None) == 1
        dummy = torch_cuda.zeros(4, dtype=torch_cuda.int32, device="cuda")
        st = lib.fsr_exclusive_scan_noalloc(h, dummy.data_ptr(), dummy.data_ptr(), (1 << 30) + 1,
                                            None, None)
        assert st == 3, st
        assert b"FSR_NOALLOC_SCAN_MAX_ELEMENTS" in lib.fsr_last_error()
        assert lib.fsr_exclusive_scan_noalloc(None, dummy.data_ptr(), dummy.data_ptr(), 4, None,
                                              None) == 1
    finally:
        ctx.close()
