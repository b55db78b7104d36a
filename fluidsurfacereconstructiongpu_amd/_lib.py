"""ctypes binding of libfsr.so (the C-ABI declared in include/fsr.h).

The library is built in-tree (``__graft_entry__.build()`` or ``make -C
fluidsurfacereconstructiongpu_amd/csrc``).  Loading fails loudly when it is
missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_uint32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfsr.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "fsr.h")


class Float3(ctypes.Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]


class UInt3(ctypes.Structure):
    _fields_ = [("x", c_uint32), ("y", c_uint32), ("z", c_uint32)]


class SimParam(ctypes.Structure):
    """fsr_sim_param == the reference's SimParam (GPU/Defines.h:18-32)."""

    _fields_ = [
        ("particle_radius", c_float),
        ("particle_mass", c_float),
        ("smoothing_radius", c_float),
        ("smoothing_radius_sq", c_float),
        ("smoothing_radius_inv", c_float),
        ("lambda_for_smoothed", c_float),
        ("anisotropic_radius", c_float),
        ("iso_value", c_float),
        ("sc_sp_grid_res_ratio", c_uint32),
        ("expand_extent", c_uint32),
        ("min_num_neighbors", c_uint32),
        ("spatial_cell_size_scale", c_float),
    ]


class GridInfo(ctypes.Structure):
    """fsr_grid_info == the reference's GridInfo (GPU/Defines.h:62-68)."""

    _fields_ = [("min_pos", Float3), ("max_pos", Float3), ("cell_size", c_float),
                ("resolution", UInt3)]

    @property
    def res(self):
        return (self.resolution.x, self.resolution.y, self.resolution.z)

    @property
    def count(self):
        r = self.res
        return r[0] * r[1] * r[2]


class AnisoParam(ctypes.Structure):
    _fields_ = [("kr", c_float), ("ks", c_float), ("kn", c_float), ("reserved", c_uint32)]


class XyzHeader(ctypes.Structure):
    """fsr_xyz_header: the five header lines of a .xyz frame."""

    _fields_ = [("aabb_min", Float3), ("aabb_max", Float3), ("kernel_radius", c_float),
                ("particle_radius", c_float), ("particle_mass", c_float)]


class FrameStats(ctypes.Structure):
    _fields_ = [
        ("num_particles", c_uint32),
        ("num_out_of_grid", c_uint32),
        ("num_surface_particles", c_uint32),
        ("num_involved_particles", c_uint32),
        ("num_band_vertices", c_uint32),
        ("num_valid_cubes", c_uint32),
        ("num_mesh_vertices", c_uint32),
        ("overflow", c_uint32),
        ("band_exceptions", c_uint32),
        ("num_unique_vertices", c_uint32),
        ("reserved", c_uint32 * 6),
    ]


STATS_WORDS = ctypes.sizeof(FrameStats) // 4


class FrameDesc(ctypes.Structure):
    """fsr_frame_desc."""

    _fields_ = [
        ("src_pos", c_void_p), ("src_dens", c_void_p), ("pos", c_void_p), ("dens", c_void_p),
        ("num_particles", c_uint32), ("tri_capacity", c_uint32), ("sp_grid", GridInfo),
        ("sc_grid", GridInfo), ("sim", SimParam), ("phi", c_void_p), ("tri_pos", c_void_p),
        ("tri_nrm", c_void_p), ("stats", c_void_p),
    ]

STAGES = ["hash", "surface", "band", "prep", "field", "mc"]

METHOD_ZB05 = 0
METHOD_AKINCI12 = 1
METHOD_YU13 = 2

BUF = {
    "cell_keys": 0, "permutation": 1, "cell_ranges": 2, "occupancy": 3,
    "surface_flags": 4, "band_bits": 5, "band_index": 6, "cube_nverts": 7,
    "valid_cubes": 8, "involve_counts": 9, "mean_pos": 10, "smooth_pos": 11,
    "g_matrices": 12, "singular": 13, "involved_index": 14, "field_work": 15,
    "color_work": 16, "aniso_work": 17,
}

STATUS = {0: "FSR_OK", 1: "FSR_ERR_INVALID_ARG", 2: "FSR_ERR_HIP", 3: "FSR_ERR_CAPACITY",
          4: "FSR_ERR_OUT_OF_MEMORY", 5: "FSR_ERR_NO_DEVICE", 6: "FSR_ERR_UNSUPPORTED"}


class FsrError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libfsr.so once.  torch is imported first so that one HIP runtime
    serves both."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401
    except Exception:  # pragma: no cover - torch is part of the image
        pass
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with __graft_entry__.build() "
            "(there is no fallback implementation)")
    lib = ctypes.CDLL(path)
    st = c_int
    lib.fsr_create.argtypes = [POINTER(c_void_p), c_int]
    lib.fsr_create.restype = st
    lib.fsr_destroy.argtypes = [c_void_p]
    lib.fsr_destroy.restype = st
    lib.fsr_last_error.argtypes = []
    lib.fsr_last_error.restype = c_char_p
    lib.fsr_status_string.argtypes = [c_int]
    lib.fsr_status_string.restype = c_char_p
    lib.fsr_version.argtypes = []
    lib.fsr_version.restype = c_char_p
    lib.fsr_grid_from_aabb.argtypes = [Float3, Float3, c_float, c_float, c_uint32,
                                       POINTER(GridInfo), POINTER(GridInfo)]
    lib.fsr_grid_from_aabb.restype = st
    lib.fsr_reconstruct_frame.argtypes = [
        c_void_p, c_int, POINTER(SimParam), POINTER(AnisoParam), POINTER(GridInfo),
        POINTER(GridInfo), c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p,
        c_uint32, c_void_p, c_void_p]
    lib.fsr_reconstruct_frame.restype = st
    lib.fsr_reconstruct_frame_indexed.argtypes = [
        c_void_p, c_int, POINTER(SimParam), POINTER(AnisoParam), POINTER(GridInfo),
        POINTER(GridInfo), c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p,
        c_uint32, c_void_p, c_uint32, c_void_p, c_void_p]
    lib.fsr_reconstruct_frame_indexed.restype = st
    lib.fsr_frame_buffer.argtypes = [c_void_p, c_int, POINTER(c_void_p), POINTER(c_uint64)]
    lib.fsr_frame_buffer.restype = st
    P = c_void_p
    stage_sigs = {
        "fsr_bind_textures": [P, P, P, P, P],
        "fsr_spatial_grid_building": [P, P, P, c_uint32, P, P, P, P],
        "fsr_extraction_of_surface_particles": [P, P, P, P, c_uint32, P, P, P],
        "fsr_extraction_of_surface_and_involve_particles": [P, P, P, P, P, c_uint32, P, P, P],
        "fsr_extraction_of_surface_particles_for_akinci": [P, P, P, P, c_uint32, P, P, P, P],
        "fsr_estimation_of_surface_vertices": [P, P, P, P, P, c_uint32, P, P, P],
        "fsr_thrust_exclusive_prefix_sum_scan": [P, P, P, c_uint32, POINTER(c_uint32), P],
        "fsr_compact_surface_vertex": [P, P, P, P, c_uint64, P],
        "fsr_update_scalar_grid_values_compacted": [P, P, c_uint32, P, P, c_uint32, P, P, P, P, P],
        "fsr_detect_valid_surface_cubes": [P, P, c_uint32, P, P, P, P, P, P, P],
        "fsr_detection_of_valid_surface_cubes_for_akinci": [P, P, c_uint32, P, P, P, P, P, P],
        "fsr_compact_valid_surface_cubes": [P, P, P, P, c_uint32, P],
        "fsr_generate_triangles": [P, P, P, c_uint32, P, P, P, P, P, c_uint32, P, P],
        "fsr_calculation_of_mean_and_smooth_particles": [P, P, P, c_uint32, P, P, P, P, P],
        "fsr_compactation_of_surface_vertices_and_particles": [P, P, P, c_uint64, P, P, P,
                                                               c_uint64, P, P, P],
        "fsr_calculation_of_transform_matrices_for_particles": [P, P, P, P, P, P, c_uint32, P, P,
                                                                c_uint32, P, P],
        "fsr_computation_of_scalar_field_grid": [P, P, c_uint32, P, P, P, P, P, c_uint32, P, P, P,
                                                 P, P, P],
        "fsr_estimate_surface_vertices": [P, P, P, P, c_uint32, P, P, P],
        "fsr_update_scalar_grid_values_tc01": [P, P, c_uint32, P, P, c_uint32, P, P, P, P, P],
    }
    for name, sig in stage_sigs.items():
        f = getattr(lib, name)
        f.argtypes = sig
        f.restype = st
    lib.fsr_exclusive_scan_noalloc.argtypes = [P, P, P, c_uint32, P, P]
    lib.fsr_exclusive_scan_noalloc.restype = st
    lib.fsr_batch_create.argtypes = [c_void_p, c_uint32, POINTER(c_void_p)]
    lib.fsr_batch_create.restype = st
    lib.fsr_batch_destroy.argtypes = [c_void_p]
    lib.fsr_batch_destroy.restype = st
    lib.fsr_batch_run.argtypes = [c_void_p, c_int, POINTER(AnisoParam), POINTER(FrameDesc),
                                  c_uint32, c_void_p, c_void_p]
    lib.fsr_batch_run.restype = st
    lib.fsr_autotune.argtypes = [c_void_p, c_int, POINTER(AnisoParam), POINTER(FrameDesc),
                                 c_uint32, POINTER(c_uint32), c_void_p]
    lib.fsr_autotune.restype = st
    lib.fsr_batch_autotune.argtypes = [c_void_p, c_int, POINTER(AnisoParam), POINTER(FrameDesc),
                                       c_uint32, c_uint32, POINTER(c_uint32), c_void_p]
    lib.fsr_batch_autotune.restype = st
    lib.fsr_set_launch_shape.argtypes = [c_void_p, POINTER(c_uint32)]
    lib.fsr_set_launch_shape.restype = st
    lib.fsr_set_profiling.argtypes = [c_void_p, c_uint32]
    lib.fsr_set_profiling.restype = st
    lib.fsr_stage_times.argtypes = [c_void_p, POINTER(c_float), c_uint32, POINTER(c_uint32)]
    lib.fsr_stage_times.restype = st
    lib.fsr_xyz_info.argtypes = [c_char_p, POINTER(XyzHeader), POINTER(c_uint32)]
    lib.fsr_xyz_info.restype = st
    lib.fsr_xyz_read.argtypes = [c_char_p, POINTER(XyzHeader), c_void_p, c_void_p, c_uint32,
                                 POINTER(c_uint32)]
    lib.fsr_xyz_read.restype = st
    lib.fsr_obj_write_soup.argtypes = [c_char_p, c_void_p, c_void_p, c_uint32]
    lib.fsr_obj_write_soup.restype = st
    lib.fsr_obj_write_indexed.argtypes = [c_char_p, c_void_p, c_void_p, c_uint32, c_void_p,
                                          c_uint32]
    lib.fsr_obj_write_indexed.restype = st
    lib.fsr_sim_param_for.argtypes = [c_int, POINTER(XyzHeader), POINTER(SimParam)]
    lib.fsr_sim_param_for.restype = st
    lib.fsr_vis_write.argtypes = [c_void_p, c_char_p, c_int, POINTER(GridInfo), POINTER(GridInfo),
                                  POINTER(SimParam), c_void_p, c_uint32, POINTER(FrameStats),
                                  c_char_p]
    lib.fsr_vis_write.restype = st
    lib.fsr_set_surface_mode.argtypes = [c_void_p, c_int]
    lib.fsr_set_surface_mode.restype = st
    lib.fsr_set_adjoint.argtypes = [c_void_p, c_int]
    lib.fsr_set_adjoint.restype = st
    lib.fsr_mc_vertex_backward.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
    lib.fsr_mc_vertex_backward.restype = st
    lib.fsr_field_zb05_backward.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
    lib.fsr_field_zb05_backward.restype = st
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        msg = _lib.fsr_last_error().decode() if _lib is not None else ""
        raise FsrError(status, msg)


def header_symbols(path: str = HEADER_PATH) -> list[str]:
    """Names of the functions declared in include/fsr.h."""
    import re

    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fsr_[a-z0-9_]+)\s*\(", text)))
