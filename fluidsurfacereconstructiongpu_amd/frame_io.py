"""Python side of the reference's frame files (.xyz in, .obj out) through the
host-only entry points of libfsr (fsr_xyz_info / fsr_xyz_read /
fsr_obj_write_soup / fsr_obj_write_indexed / fsr_sim_param_for).  The formats
are those of FSR/data_format.txt and GPU/ReconstructorGPUOurs.cpp:20-100."""
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
