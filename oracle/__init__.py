"""Python binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as part of the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_MC = os.path.join(_HERE, "_ref", "libref_mctables.so")

u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OGrid(ctypes.Structure):
    _fields_ = [("minx", ctypes.c_float), ("miny", ctypes.c_float), ("minz", ctypes.c_float),
                ("cell", ctypes.c_float), ("rx", ctypes.c_uint32), ("ry", ctypes.c_uint32),
                ("rz", ctypes.c_uint32)]


def ogrid(g) -> OGrid:
    return OGrid(float(g.min[0]), float(g.min[1]), float(g.min[2]), float(g.cell),
                 int(g.res[0]), int(g.res[1]), int(g.res[2]))


_lib = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", _HERE, "all"], check=True, stdout=subprocess.DEVNULL)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    lib.orc_num_threads.restype = ctypes.c_int
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_hash_sort.argtypes = [f32p, f32p, ctypes.c_uint32, OGrid, u32p, u32p, f32p, f32p,
                                  u32p, f32p]
    lib.orc_surface_occupancy.argtypes = [f32p, u32p, OGrid, u32p, ctypes.c_uint32, u8p, u32p,
                                          ctypes.c_void_p]
    lib.orc_surface_colorfield.argtypes = [f32p, f32p, ctypes.c_uint32, u32p, OGrid,
                                           ctypes.c_float, ctypes.c_float, u32p, f64p, u32p]
    lib.orc_band.argtypes = [f32p, u32p, ctypes.c_uint32, OGrid, ctypes.c_float, u8p]
    lib.orc_compact.argtypes = [u8p, ctypes.c_uint64, u32p]
    lib.orc_compact.restype = ctypes.c_uint32
    lib.orc_field_zb05.argtypes = [u32p, ctypes.c_uint32, f32p, u32p, OGrid, OGrid,
                                   ctypes.c_float, ctypes.c_float, f64p, f64p]
    lib.orc_mc.argtypes = [u32p, ctypes.c_uint32, u8p, f64p, OGrid, ctypes.c_float, ctypes.c_int,
                           u32p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                           ctypes.c_void_p]
    lib.orc_mc.restype = ctypes.c_uint64
    lib.orc_frame.argtypes = [ctypes.c_int, f32p, f32p, ctypes.c_uint32, OGrid, OGrid,
                              ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float, ctypes.c_void_p, u64p]
    lib.orc_surface_radius_count.argtypes = [f32p, ctypes.c_uint32, u32p, OGrid, ctypes.c_float,
                                             u32p, u32p]
    lib.orc_mc_tri_packed.argtypes = [ctypes.c_uint32]
    lib.orc_mc_tri_packed.restype = ctypes.c_uint64
    lib.orc_smooth.argtypes = [f32p, ctypes.c_uint32, u32p, OGrid, ctypes.c_float, ctypes.c_float,
                               f64p, f64p]
    lib.orc_eigen3.argtypes = [f64p, f64p, f64p]
    lib.orc_aniso_matrices.argtypes = [f32p, f64p, f32p, ctypes.c_uint32, u32p, u32p, OGrid,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_uint32, f64p, f64p, u32p, u32p]
    lib.orc_field_aniso.argtypes = [u32p, ctypes.c_uint32, f64p, f32p, u32p, u32p, u32p, f64p,
                                    OGrid, OGrid, ctypes.c_float, ctypes.c_void_p, f64p,
                                    ctypes.c_void_p]
    lib.orc_estimate_surface_vertices.argtypes = [f32p, OGrid, ctypes.c_uint32, ctypes.c_uint32,
                                                  OGrid, u32p]
    lib.orc_field_tc01.argtypes = [u32p, ctypes.c_uint32, f32p, u32p, OGrid, OGrid,
                                   ctypes.c_float, f64p]
    lib.orc_mc_indexed.argtypes = [u32p, ctypes.c_uint32, u8p, f64p, OGrid, ctypes.c_float,
                                   ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, u64p]
    lib.orc_mc_indexed.restype = ctypes.c_uint64
    lib.orc_mc_vertex_backward.argtypes = [u32p, ctypes.c_uint32, u8p, f64p, OGrid,
                                           ctypes.c_float, ctypes.c_int, f64p, ctypes.c_uint64,
                                           f64p]
    lib.orc_field_zb05_backward.argtypes = [u32p, ctypes.c_uint32, f32p, ctypes.c_uint32, u32p,
                                            OGrid, OGrid, ctypes.c_float, f64p, f64p]
    lib.orc_adj_create.restype = ctypes.c_void_p
    lib.orc_adj_create.argtypes = [ctypes.c_int, f32p, f32p, ctypes.c_uint32, OGrid, OGrid,
                                   ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double]
    lib.orc_adj_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_adj_info.argtypes = [ctypes.c_void_p, u64p]
    lib.orc_adj_state.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p]
    lib.orc_adj_loss.argtypes = [ctypes.c_void_p]
    lib.orc_adj_loss.restype = ctypes.c_double
    lib.orc_adj_upstream.argtypes = [ctypes.c_void_p, f64p, f64p]
    lib.orc_adj_grad.argtypes = [ctypes.c_void_p, f64p]
    lib.orc_adj_fd.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_double]
    lib.orc_adj_fd.restype = ctypes.c_double
    _lib = lib
    return lib


def smooth(pos, ranges, sp, radius, lam):
    lib = load()
    n = pos.shape[0]
    mean = np.empty((n, 3), np.float64)
    sm = np.empty((n, 3), np.float64)
    lib.orc_smooth(np.ascontiguousarray(pos, np.float32).ravel(), n,
                   np.ascontiguousarray(ranges).ravel(), ogrid(sp), radius, lam, mean.ravel(),
                   sm.ravel())
    return mean, sm


def eigen3(a):
    lib = load()
    d = np.empty(3, np.float64)
    v = np.empty(9, np.float64)
    lib.orc_eigen3(np.ascontiguousarray(a, np.float64).ravel(), d, v)
    return d, v.reshape(3, 3)


def involved_list(keys, ranges, involve):
    """Sorted ids of the particles of involved cells, in cell order, and the
    per-cell offsets (the reference's involve scan)."""
    off = np.zeros(involve.shape[0], np.uint32)
    off[1:] = np.cumsum(involve[:-1], dtype=np.uint64).astype(np.uint32)
    C = involve.shape[0]
    valid = keys < C
    sel = np.zeros(keys.shape[0], bool)
    sel[valid] = involve[keys[valid]] > 0
    return np.nonzero(sel)[0].astype(np.uint32), off


def aniso_matrices(pos, mean, dens, index, ranges, sp, h, radius, kr=4.0, kn=1.0, ne=25):
    lib = load()
    I = index.shape[0]
    G = np.empty((I, 3, 3), np.float64)
    sig = np.empty((I, 3), np.float64)
    lo = np.empty(I, np.uint32)
    hi = np.empty(I, np.uint32)
    lib.orc_aniso_matrices(np.ascontiguousarray(pos, np.float32).ravel(),
                           np.ascontiguousarray(mean, np.float64).ravel(),
                           np.ascontiguousarray(dens, np.float32), I,
                           np.ascontiguousarray(index, np.uint32),
                           np.ascontiguousarray(ranges).ravel(), ogrid(sp), h, radius, kr, kn,
                           ne, G.ravel(), sig.ravel(), lo, hi)
    return G, sig, lo, hi


def field_aniso(index, smooth_pos, dens, ranges, involve, involve_off, G, sp, sc, mass,
                ambiguous=None):
    lib = load()
    B = index.shape[0]
    phi = np.full(sc.count, np.frombuffer(np.uint32(0x64646464).tobytes(), np.float32)[0],
                  np.float64)
    flag = np.zeros(B, np.uint8)
    amb = None if ambiguous is None else np.ascontiguousarray(ambiguous, np.uint8)
    lib.orc_field_aniso(np.ascontiguousarray(index, np.uint32), B,
                        np.ascontiguousarray(smooth_pos, np.float64).ravel(),
                        np.ascontiguousarray(dens, np.float32), np.ascontiguousarray(ranges).ravel(),
                        np.ascontiguousarray(involve, np.uint32),
                        np.ascontiguousarray(involve_off, np.uint32),
                        np.ascontiguousarray(G, np.float64).ravel(), ogrid(sp), ogrid(sc), mass,
                        amb.ctypes.data if amb is not None else None, phi,
                        flag.ctypes.data)
    return phi, flag


def threads() -> int:
    return int(load().orc_num_threads())


def set_threads(n: int) -> None:
    load().orc_set_threads(int(n))


def hash_sort(pos, dens, sp):
    lib = load()
    n = pos.shape[0]
    C = sp.count
    keys = np.empty(n, np.uint32)
    perm = np.empty(n, np.uint32)
    pos_o = np.empty((n, 3), np.float32)
    dens_o = np.empty(n, np.float32)
    ranges = np.empty(2 * C, np.uint32)
    occ = np.empty(C, np.float32)
    lib.orc_hash_sort(np.ascontiguousarray(pos, np.float32).ravel(),
                      np.ascontiguousarray(dens, np.float32), n, ogrid(sp), keys, perm,
                      pos_o.ravel(), dens_o, ranges, occ)
    return dict(keys=keys, perm=perm, pos=pos_o, dens=dens_o, ranges=ranges.reshape(C, 2),
                occ=occ)


def surface_occupancy(occ, ranges, sp, keys, involve=False):
    lib = load()
    C = sp.count
    n = keys.shape[0]
    surf = np.empty(C, np.uint8)
    flags = np.empty(n, np.uint32)
    inv = np.empty(C, np.uint32) if involve else None
    lib.orc_surface_occupancy(occ, np.ascontiguousarray(ranges).ravel(), ogrid(sp), keys, n, surf,
                              flags, inv.ctypes.data if inv is not None else None)
    return surf, flags, inv


def surface_radius_count(pos, ranges, sp, h):
    """(flags, counts) of the radius-count surface test."""
    lib = load()
    n = pos.shape[0]
    flags = np.empty(n, np.uint32)
    cnt = np.empty(n, np.uint32)
    lib.orc_surface_radius_count(np.ascontiguousarray(pos, np.float32).ravel(), n,
                                 np.ascontiguousarray(ranges).ravel(), ogrid(sp), h, flags, cnt)
    return flags, cnt


def surface_colorfield(pos, dens, ranges, sp, h, mass):
    lib = load()
    n = pos.shape[0]
    flags = np.empty(n, np.uint32)
    g = np.empty(n, np.float64)
    cnt = np.empty(n, np.uint32)
    lib.orc_surface_colorfield(np.ascontiguousarray(pos, np.float32).ravel(),
                               np.ascontiguousarray(dens, np.float32), n,
                               np.ascontiguousarray(ranges).ravel(), ogrid(sp), h, mass, flags,
                               g, cnt)
    return flags, g, cnt


def band(pos, flags, sc, h):
    lib = load()
    out = np.empty(sc.count, np.uint8)
    lib.orc_band(np.ascontiguousarray(pos, np.float32).ravel(),
                 np.ascontiguousarray(flags, np.uint32), pos.shape[0], ogrid(sc), h, out)
    return out


def compact(band_u8):
    lib = load()
    idx = np.empty(band_u8.shape[0], np.uint32)
    b = lib.orc_compact(band_u8, band_u8.shape[0], idx)
    return idx[:b].copy()


def field_zb05(index, pos, ranges, sp, sc, radius_sq, particle_radius, fill=0.0):
    lib = load()
    phi = np.full(sc.count, fill, np.float64)
    wsum = np.empty(index.shape[0], np.float64)
    lib.orc_field_zb05(np.ascontiguousarray(index, np.uint32), index.shape[0],
                       np.ascontiguousarray(pos, np.float32).ravel(),
                       np.ascontiguousarray(ranges).ravel(), ogrid(sp), ogrid(sc), radius_sq,
                       particle_radius, phi, wsum)
    return phi, wsum


def mc(index, band_u8, phi, sc, iso, cpu=False, with_mesh=True, with_len=False):
    """(nverts, soup positions, soup normals[, normal lengths before the
    normalisation when with_len])."""
    lib = load()
    index = np.ascontiguousarray(index, np.uint32)
    B = index.shape[0]
    nverts = np.empty(B, np.uint32)
    phi = np.ascontiguousarray(phi, np.float64)
    total = lib.orc_mc(index, B, band_u8, phi, ogrid(sc), iso, int(cpu), nverts, None, None, 0,
                       None)
    pos = nrm = nlen = None
    if with_mesh:
        pos = np.empty((total, 3), np.float64)
        nrm = np.empty((total, 3), np.float64)
        nlen = np.empty(total, np.float64)
        lib.orc_mc(index, B, band_u8, phi, ogrid(sc), iso, int(cpu), nverts, pos.ctypes.data,
                   nrm.ctypes.data, total, nlen.ctypes.data)
    if with_len:
        return nverts, pos, nrm, nlen
    return nverts, pos, nrm


def frame(method, fr, want_phi=False, radius_count=False):
    """Whole frame on copies of the frame's arrays; returns (counts, phi|None).
    radius_count: surface particles by the radius-count test."""
    lib = load()
    pos = np.ascontiguousarray(fr.pos, np.float32).copy()
    dens = np.ascontiguousarray(fr.dens, np.float32).copy()
    phi = np.empty(fr.sc.count, np.float32) if want_phi else None
    counts = np.zeros(6, np.uint64)
    h = np.float32(fr.h)
    lib.orc_frame({"zb05": 0, "akinci12": 1}[method] | (2 if radius_count else 0), pos.ravel(),
                  dens, pos.shape[0],
                  ogrid(fr.sp), ogrid(fr.sc), float(h), float(h * h),
                  float(np.float32(fr.particle_radius)), float(np.float32(fr.mass)),
                  float(np.float32(fr.iso)), phi.ctypes.data if want_phi else None, counts)
    keys = ["surface_particles", "band_vertices", "valid_cubes", "mesh_vertices", "out_of_grid"]
    return {k: int(counts[i]) for i, k in enumerate(keys)}, phi


def mc_tri_packed():
    lib = load()
    return np.array([lib.orc_mc_tri_packed(c) for c in range(256)], dtype=np.uint64)


def reference_mc_tables():
    """The reference's own MC tables (oracle/_ref, built from the reference's
    GPU/MarchingCubesHelper.cpp), or None when that build is absent."""
    if not os.path.exists(REF_MC):
        return None
    lib = ctypes.CDLL(REF_MC)
    for f in ("ref_mc_edge_flags", "ref_mc_tri_table", "ref_mc_num_vertices",
              "ref_mc_edge_vertices", "ref_mc_corner_offsets"):
        getattr(lib, f).restype = ctypes.c_void_p

    def arr(addr, ctype, n):
        return np.ctypeslib.as_array((ctype * n).from_address(addr)).copy()

    return dict(
        edge_flags=arr(lib.ref_mc_edge_flags(), ctypes.c_uint32, 256),
        tri=arr(lib.ref_mc_tri_table(), ctypes.c_int32, 256 * 16).reshape(256, 16),
        num_vertices=arr(lib.ref_mc_num_vertices(), ctypes.c_uint32, 256),
        edge_vertices=arr(lib.ref_mc_edge_vertices(), ctypes.c_uint32, 24).reshape(12, 2),
        corner_offsets=arr(lib.ref_mc_corner_offsets(), ctypes.c_int32, 24).reshape(8, 3),
    )


def estimate_surface_vertices(dens_grid, dg, ratio, ext, sg):
    """u32[V] ones on the vertices estimateSurfaceVertices marks."""
    lib = load()
    out = np.zeros(sg.count, np.uint32)
    lib.orc_estimate_surface_vertices(np.ascontiguousarray(dens_grid, np.float32), ogrid(dg),
                                      int(ratio), int(ext), ogrid(sg), out)
    return out


def field_tc01(index, pos_sorted, ranges, sp, sc, radius_sq, fill=0.0):
    lib = load()
    phi = np.full(sc.count, fill, np.float64)
    index = np.ascontiguousarray(index, np.uint32)
    lib.orc_field_tc01(index, index.shape[0], np.ascontiguousarray(pos_sorted, np.float32).ravel(),
                       np.ascontiguousarray(ranges).ravel(), ogrid(sp), ogrid(sc), radius_sq, phi)
    return phi


def mc_indexed(index, band_u8, phi, sc, iso, cpu=False):
    """Deduplicated mesh over `phi`: (vertices (U, 3), normals (U, 3), indices (3T,))."""
    lib = load()
    index = np.ascontiguousarray(index, np.uint32)
    B = index.shape[0]
    phi = np.ascontiguousarray(phi, np.float64)
    band_u8 = np.ascontiguousarray(band_u8, np.uint8)
    nidx = np.zeros(1, np.uint64)
    U = int(lib.orc_mc_indexed(index, B, band_u8, phi, ogrid(sc), iso, int(cpu), None, None, 0,
                               None, 0, nidx))
    T3 = int(nidx[0])
    pos = np.empty((U, 3), np.float64)
    nrm = np.empty((U, 3), np.float64)
    idx = np.empty(T3, np.uint32)
    lib.orc_mc_indexed(index, B, band_u8, phi, ogrid(sc), iso, int(cpu), pos.ctypes.data,
                       nrm.ctypes.data, U, idx.ctypes.data, T3, nidx)
    return pos, nrm, idx


def mc_vertex_backward(index, band_u8, phi, sc, iso, g_tri, cpu=False, g_phi=None):
    """g_phi (float64[V], accumulated) += the marching-cubes vertex term for the
    soup over `phi` (soup order of the device emitter)."""
    lib = load()
    index = np.ascontiguousarray(index, np.uint32)
    out = np.zeros(sc.count, np.float64) if g_phi is None else np.array(g_phi, np.float64)
    g = np.ascontiguousarray(g_tri, np.float64).ravel()
    lib.orc_mc_vertex_backward(index, index.shape[0], np.ascontiguousarray(band_u8, np.uint8),
                               np.ascontiguousarray(phi, np.float64), ogrid(sc), iso, int(cpu),
                               g, g.shape[0] // 3, out)
    return out


def field_zb05_backward(index, pos_sorted, ranges, sp, sc, radius_sq, g_phi):
    """dL/dp (float64[n, 3], hash-sorted order) from g_phi (float64[V])."""
    lib = load()
    index = np.ascontiguousarray(index, np.uint32)
    n = pos_sorted.shape[0]
    out = np.empty((n, 3), np.float64)
    lib.orc_field_zb05_backward(index, index.shape[0],
                                np.ascontiguousarray(pos_sorted, np.float32).ravel(), n,
                                np.ascontiguousarray(ranges).ravel(), ogrid(sp), ogrid(sc),
                                radius_sq, np.ascontiguousarray(g_phi, np.float64), out.ravel())
    return out


class AdjointWorkload:
    """FP64 adjoint workload on a frame: field + soup with frozen topology, the
    sphere loss, its analytic gradient and incremental central differences."""

    def __init__(self, method, fr, sphere_radius=0.3, w_grid=1.0, w_soup=1.0):
        self.lib = load()
        h = np.float32(fr.h)
        pos = np.ascontiguousarray(fr.pos, np.float32)
        dens = np.ascontiguousarray(fr.dens, np.float32)
        self.n = pos.shape[0]
        self.ptr = self.lib.orc_adj_create(
            {"zb05": 0, "akinci12": 1}[method], pos.ravel(), dens, self.n, ogrid(fr.sp),
            ogrid(fr.sc), float(h), float(h * h), float(np.float32(fr.particle_radius)),
            float(np.float32(fr.mass)), float(np.float32(fr.iso)), float(sphere_radius),
            float(w_grid), float(w_soup))
        info = np.zeros(6, np.uint64)
        self.lib.orc_adj_info(self.ptr, info)
        self.B, self.K, self.V = int(info[1]), int(info[2]), int(info[3])

    def close(self):
        if self.ptr:
            self.lib.orc_adj_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def state(self):
        phi = np.empty(self.V, np.float64)
        soup = np.empty((self.K, 3), np.float64)
        index = np.empty(self.B, np.uint32)
        perm = np.empty(self.n, np.uint32)
        self.lib.orc_adj_state(self.ptr, phi.ctypes.data, soup.ctypes.data, index.ctypes.data,
                               perm.ctypes.data)
        return dict(phi=phi, soup=soup, band_index=index, perm=perm)

    def loss(self) -> float:
        return float(self.lib.orc_adj_loss(self.ptr))

    def upstream(self):
        g_phi = np.empty(self.V, np.float64)
        g_soup = np.empty((self.K, 3), np.float64)
        self.lib.orc_adj_upstream(self.ptr, g_phi, g_soup.ravel())
        return g_phi, g_soup

    def grad(self):
        g = np.empty((self.n, 3), np.float64)
        self.lib.orc_adj_grad(self.ptr, g.ravel())
        return g

    def fd(self, j: int, axis: int, eps: float = 1e-6) -> float:
        return float(self.lib.orc_adj_fd(self.ptr, int(j), int(axis), float(eps)))


def frame_yu13(fr, aniso):
    """Whole anisotropic (Yu13) frame through the oracle's stage functions:
    hashing, surface + involved cells, band, smoothing, involved list,
    covariance -> G, field, marching cubes.  Returns the frame's counts."""
    f = np.float32
    h = f(fr.h)
    r = f(h * f(2.0))
    hs = hash_sort(fr.pos, fr.dens, fr.sp)
    _, flags, involve = surface_occupancy(hs["occ"], hs["ranges"], fr.sp, hs["keys"],
                                          involve=True)
    band_u8 = band(hs["pos"], flags, fr.sc, h)
    index = compact(band_u8)
    mean, sm = smooth(hs["pos"], hs["ranges"], fr.sp, r, f(aniso["lam"]))
    inv_idx, inv_off = involved_list(hs["keys"], hs["ranges"], involve)
    G, _, _, _ = aniso_matrices(hs["pos"], mean, hs["dens"], inv_idx, hs["ranges"], fr.sp, h, r,
                                aniso["kr"], aniso["kn"], aniso["ne"])
    phi, _ = field_aniso(index, sm, hs["dens"], hs["ranges"], involve, inv_off, G, fr.sp, fr.sc,
                         f(fr.mass))
    nv, _, _ = mc(index, band_u8, phi, fr.sc, f(fr.iso), with_mesh=False)
    return {"surface_particles": int(flags.sum()), "involved_particles": int(inv_idx.shape[0]),
            "band_vertices": int(index.shape[0]), "valid_cubes": int((nv > 0).sum()),
            "mesh_vertices": int(nv.sum()), "out_of_grid": 0}
