"""The C-ABI library loads, exports every function include/fsr.h declares, and
its host-only entry points behave without a GPU."""
import ctypes

import numpy as np
import pytest

from fluidsurfacereconstructiongpu_amd import _lib
from fluidsurfacereconstructiongpu_amd import datasets as ds
from tests.conftest import gpu_available


def test_header_symbols_exported():
    lib = _lib.load()
    names = _lib.header_symbols()
    assert "fsr_reconstruct_frame" in names and len(names) >= 8
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_reference_layouts():
    # SimParam 12 x 4 bytes, GridInfo 40 bytes (GPU/Defines.h:18-68)
    assert ctypes.sizeof(_lib.SimParam) == 48
    assert ctypes.sizeof(_lib.GridInfo) == 40
    assert ctypes.sizeof(_lib.FrameStats) == 64


@pytest.mark.parametrize("h,scale,ratio", [(0.02, 1.0, 2), (0.02, 1.3, 3), (0.036, 1.0, 2)])
def test_grid_rule_matches_fp32_restatement(h, scale, ratio):
    lib = _lib.load()
    lo = np.array([-0.04, -0.0371, 0.0123], np.float32)
    hi = np.array([3.04, 2.04, 1.54], np.float32)
    sp, sc = _lib.GridInfo(), _lib.GridInfo()
    _lib.check(lib.fsr_grid_from_aabb(_lib.Float3(*map(float, lo)), _lib.Float3(*map(float, hi)),
                                      h, scale, ratio, ctypes.byref(sp), ctypes.byref(sc)))
    g = ds.spatial_grid(lo, hi, h, scale)
    s = ds.scalar_grid_ratio(g, ratio)
    assert sp.res == g.res and sc.res == s.res
    assert np.float32(sp.cell_size) == g.cell and np.float32(sc.cell_size) == s.cell
    assert np.float32(sp.min_pos.x) == g.min[0] and np.float32(sc.max_pos.z) == s.max[2]


def test_error_paths_without_device_work():
    lib = _lib.load()
    assert lib.fsr_grid_from_aabb(_lib.Float3(0, 0, 0), _lib.Float3(1, 1, 1), -1.0, 1.0, 2,
                                  None, None) == 1
    assert b"invalid" in lib.fsr_last_error()
    assert lib.fsr_status_string(3) == b"output capacity exceeded"
    assert lib.fsr_version().startswith(b"fsr")
    h = ctypes.c_void_p()
    st = lib.fsr_create(ctypes.byref(h), 0)
    if gpu_available():
        assert st == 0
        lib.fsr_destroy(h)
    else:
        assert st == 5  # FSR_ERR_NO_DEVICE
    assert lib.fsr_reconstruct_frame(None, 0, None, None, None, None, None, None, 0, None, None,
                                     None, 0, None, None) == 1


def test_adjoint_entry_points_reject_missing_state():
    lib = _lib.load()
    assert lib.fsr_set_adjoint(None, 1) == 1
    assert lib.fsr_mc_vertex_backward(None, None, None, None) == 1
    assert lib.fsr_field_zb05_backward(None, None, None, None) == 1
    assert b"null context" in lib.fsr_last_error()


def test_vis_write_rejects_missing_arguments():
    lib = _lib.load()
    g = _lib.GridInfo()
    sim = _lib.SimParam()
    st = _lib.FrameStats()
    assert lib.fsr_vis_write(None, b"/tmp/x.vis", 0, ctypes.byref(g), ctypes.byref(g),
                             ctypes.byref(sim), None, 0, ctypes.byref(st), b"x.obj") == 1
    assert b"null argument" in lib.fsr_last_error()
