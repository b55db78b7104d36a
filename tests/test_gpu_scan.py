"""The exclusive scan behind every prefix sum of the pipeline (two passes:
segment partials, then each segment rescanned from the sum of the partials
before it), through the reference's scan launcher
(launchThrustExclusivePrefixSumScan, ReconstructionCUDA.cu:1154-1172:
exclusive scan plus the total).  Checked against numpy's cumulative sum (u32,
wrapping) on tile boundaries (4096 elements per tile), more tiles than
workgroups (1024), in-place and unaligned buffers, and many launches in a row
on one context."""
import ctypes

import numpy as np
import pytest

from fluidsurfacereconstructiongpu_amd import _lib
from fluidsurfacereconstructiongpu_amd.reconstructor import Reconstructor

pytestmark = pytest.mark.gpu


def expected(x):
    ex = np.zeros_like(x)
    if x.size:
        c = np.cumsum(x, dtype=np.uint32)
        ex[1:] = c[:-1]
        return ex, int(c[-1])
    return ex, 0


class Scan:
    def __init__(self, torch):
        self.t = torch
        self.rec = Reconstructor("zb05")
        self.lib = self.rec.ctx.lib
        self.ctx = self.rec.ctx.handle.value
        self.dev = self.rec.device

    def __call__(self, x, in_place=False, offset=0):
        t = self.t
        n = x.size
        buf = t.zeros(n + offset + 1, dtype=t.int32, device=self.dev)
        buf[offset:offset + n] = t.from_numpy(x.view(np.int32)).to(self.dev)
        out_buf = buf if in_place else t.full((n + offset + 1,), -1, dtype=t.int32,
                                              device=self.dev)
        total = ctypes.c_uint32(12345)
        elem = 4
        st = self.lib.fsr_thrust_exclusive_prefix_sum_scan(
            self.ctx, out_buf.data_ptr() + elem * offset, buf.data_ptr() + elem * offset, n,
            ctypes.byref(total), None)
        _lib.check(st)
        t.cuda.synchronize()
        got = out_buf[offset:offset + n].cpu().numpy().view(np.uint32)
        return got, total.value


@pytest.mark.parametrize("n", [1, 7, 4095, 4096, 4097, 8192, 1024 * 4096, 1024 * 4096 + 1,
                               300_001, 5_000_000])
def test_scan_sizes(torch_cuda, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1000, size=n, dtype=np.uint32)
    S = Scan(torch_cuda)
    got, tot = S(x)
    ex, etot = expected(x)
    np.testing.assert_array_equal(got, ex)
    assert tot == etot


def test_scan_empty(torch_cuda):
    S = Scan(torch_cuda)
    got, tot = S(np.zeros(0, np.uint32))
    assert got.size == 0 and tot == 0


def test_scan_in_place_unaligned_and_wrapping(torch_cuda):
    rng = np.random.default_rng(7)
    S = Scan(torch_cuda)
    x = rng.integers(0, 2 ** 32, size=777_777, dtype=np.uint64).astype(np.uint32)
    for in_place in (False, True):
        for offset in (0, 1, 3):
            got, tot = S(x, in_place=in_place, offset=offset)
            ex, etot = expected(x)
            np.testing.assert_array_equal(got, ex)
            assert tot == etot


def test_scan_many_launches(torch_cuda):
    """Consecutive launches of different sizes on one context (the partials
    buffer is shared)."""
    rng = np.random.default_rng(11)
    S = Scan(torch_cuda)
    for k in range(60):
        n = int(rng.choice([3000, 150_000, 600_000, 4096 * 3, 2_000_000]))
        x = rng.integers(0, 50, size=n, dtype=np.uint32)
        got, tot = S(x)
        ex, etot = expected(x)
        np.testing.assert_array_equal(got, ex, err_msg=f"launch {k}, n={n}")
        assert tot == etot
