"""Python side of the reference's frame files (.xyz in, .obj and .vis out)
through the entry points of libfsr (fsr_xyz_info / fsr_xyz_read /
fsr_obj_write_soup / fsr_obj_write_indexed / fsr_vis_write /
fsr_sim_param_for).  The formats are those of FSR/data_format.txt,
GPU/ReconstructorGPUOurs.cpp:20-100 and saveMiddleDataToVisFile
(GPU/ReconstructorGPUOursZB05.cpp:104-231)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

METHODS = {"zb05": _lib.METHOD_ZB05, "akinci12": _lib.METHOD_AKINCI12, "yu13": _lib.METHOD_YU13}
DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fsr_reconstruct")


def read_xyz(path: str):
    """(header dict, positions float32 (N, 3), densities float32 (N,))."""
    lib = _lib.load()
    h = _lib.XyzHeader()
    n = ctypes.c_uint32()
    _lib.check(lib.fsr_xyz_info(path.encode(), ctypes.byref(h), ctypes.byref(n)))
    pos = np.empty((n.value, 3), np.float32)
    dens = np.empty(n.value, np.float32)
    _lib.check(lib.fsr_xyz_read(path.encode(), ctypes.byref(h), pos.ctypes.data,
                                dens.ctypes.data, n.value, ctypes.byref(n)))
    head = dict(aabb_min=(h.aabb_min.x, h.aabb_min.y, h.aabb_min.z),
                aabb_max=(h.aabb_max.x, h.aabb_max.y, h.aabb_max.z),
                kernel_radius=h.kernel_radius, particle_radius=h.particle_radius,
                particle_mass=h.particle_mass)
    return head, pos[: n.value], dens[: n.value]


def write_xyz(path: str, aabb_min, aabb_max, kernel_radius, particle_radius, particle_mass,
              pos, dens) -> None:
    """Writes a .xyz frame whose numbers parse back to the same FP32 values."""
    f32 = np.float32
    with open(path, "w") as f:
        f.write("%.9g %.9g %.9g\n" % tuple(float(f32(v)) for v in aabb_min))
        f.write("%.9g %.9g %.9g\n" % tuple(float(f32(v)) for v in aabb_max))
        for v in (kernel_radius, particle_radius, particle_mass):
            f.write("%.9g\n" % float(f32(v)))
        p = np.asarray(pos, np.float32)
        d = np.asarray(dens, np.float32)
        np.savetxt(f, np.column_stack([p.astype(np.float64), d.astype(np.float64)]),
                   fmt="%.9g")


def write_obj_soup(path: str, pos, nrm=None) -> None:
    lib = _lib.load()
    p = np.ascontiguousarray(pos, np.float32)
    n = None if nrm is None else np.ascontiguousarray(nrm, np.float32)
    _lib.check(lib.fsr_obj_write_soup(path.encode(), p.ctypes.data,
                                      None if n is None else n.ctypes.data, p.shape[0]))


def write_obj_indexed(path: str, pos, nrm, indices) -> None:
    lib = _lib.load()
    p = np.ascontiguousarray(pos, np.float32)
    n = None if nrm is None else np.ascontiguousarray(nrm, np.float32)
    i = np.ascontiguousarray(indices, np.uint32)
    _lib.check(lib.fsr_obj_write_indexed(path.encode(), p.ctypes.data,
                                         None if n is None else n.ctypes.data, p.shape[0],
                                         i.ctypes.data, i.shape[0]))


def sim_param_for(method: str, header: dict) -> _lib.SimParam:
    lib = _lib.load()
    h = _lib.XyzHeader()
    h.aabb_min = _lib.Float3(*header["aabb_min"])
    h.aabb_max = _lib.Float3(*header["aabb_max"])
    h.kernel_radius = header["kernel_radius"]
    h.particle_radius = header["particle_radius"]
    h.particle_mass = header["particle_mass"]
    out = _lib.SimParam()
    _lib.check(lib.fsr_sim_param_for(METHODS[method], ctypes.byref(h), ctypes.byref(out)))
    return out


def write_vis(rec, path: str, fr, stats: dict, obj_name: str) -> None:
    """The last frame of `rec` (a Reconstructor) as the reference's .vis file."""
    from .reconstructor import grid_info
    lib = _lib.load()
    st = _lib.FrameStats()
    for name, _ in _lib.FrameStats._fields_:
        if name in stats:
            setattr(st, name, int(stats[name]))
    spg, scg = grid_info(fr.sp), grid_info(fr.sc)
    sim = rec.params(fr)
    _lib.check(lib.fsr_vis_write(rec.ctx.handle, path.encode(), METHODS[rec.method],
                                 ctypes.byref(spg), ctypes.byref(scg), ctypes.byref(sim),
                                 ctypes.c_void_p(rec.pos_sorted.data_ptr()), int(fr.n),
                                 ctypes.byref(st), obj_name.encode()))


def read_vis(path: str) -> dict:
    """Parses a .vis file (for checks): grids, radius, particles, occupied
    cells, smoothed particles, band vertices, surface particles, involved
    particles, valid cubes, mesh file name, smoothing radius."""
    with open(path) as f:
        lines = f.read().split("\n")
    it = iter(lines)

    def nums(line, t=float):
        return [t(x) for x in line.split()]

    def grid():
        return dict(min=nums(next(it)), max=nums(next(it)), res=nums(next(it), int),
                    cell=float(next(it)))

    def idx_list():
        k = int(next(it))
        return np.array(nums(next(it), int), np.int64) if k else np.zeros(0, np.int64), k

    def points():
        k = int(next(it))
        return np.array([nums(next(it)) for _ in range(k)], np.float64).reshape(k, 3)

    out = dict(sp=grid(), sc=grid(), particle_radius=float(next(it)))
    out["particles"] = points()
    out["occupied"], _ = idx_list()
    out["smoothed"] = points()
    out["band"], _ = idx_list()
    out["surface"], _ = idx_list()
    out["involved"], _ = idx_list()
    out["valid_cubes"], _ = idx_list()
    out["obj"] = next(it)
    out["smoothing_radius"] = float(next(it))
    return out
