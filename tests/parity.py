"""Shared helpers of the GPU parity tests: run a frame through libfsr.so and
compare each stage with the CPU oracle.

Tolerances (written here once, used by every GPU test):
  * hashing, sort order, cell ranges, occupancy, surface flags, band, band
    index, cube case indices of a given field: bit-exact;
  * ZB05 field: |dphi| <= 1e-5 * max(|phi|, h);
  * marching-cubes vertices of a given field, in grid-index units:
    |dv| <= 1e-5 + 2 ulp(|v|) / cell; normals |dn| <= 1e-5 + 2e-6 / L (NaN ==
    NaN), where L is the length of the interpolated normal before its
    normalisation (FP64 oracle): FP32 corner normals and lerp leave ~1e-6 of
    absolute error in it, which the normalisation divides by L;
  * cube indices of the GPU field vs the FP64 field: identical except cubes
    with a corner within 1e-5 of the iso value.
"""
from __future__ import annotations

import numpy as np

import oracle
from fluidsurfacereconstructiongpu_amd.reconstructor import Reconstructor

FIELD_RTOL = 1e-5
VERTEX_TOL = 1e-5
NORMAL_TOL = 1e-5
ISO_EXEMPT = 1e-5


def unpack_bits(words: np.ndarray, res) -> np.ndarray:
    rx, ry, rz = res
    wx = (rx + 31) // 32
    w = words.reshape(rz, ry, wx)
    bits = np.unpackbits(w.view(np.uint8).reshape(rz, ry, wx * 4), axis=2, bitorder="little")
    return bits[:, :, :rx].reshape(-1).astype(np.uint8)


def run_gpu(fr, method="zb05", rec=None):
    rec = rec or Reconstructor(method)
    res = rec.run(fr)
    return rec, res


class Stage:
    """Oracle results of every stage of one frame (ZB05 or Akinci12)."""

    def __init__(self, fr, method="zb05", radius_count=False):
        self.fr = fr
        h = np.float32(fr.h)
        self.hs = oracle.hash_sort(fr.pos, fr.dens, fr.sp)
        if radius_count:
            self.flags, self.count = oracle.surface_radius_count(
                self.hs["pos"], self.hs["ranges"], fr.sp, h)
        elif method == "akinci12":
            self.flags, self.gnorm, self.count = oracle.surface_colorfield(
                self.hs["pos"], self.hs["dens"], self.hs["ranges"], fr.sp, h,
                np.float32(fr.mass))
        else:
            self.surf_cell, self.flags, _ = oracle.surface_occupancy(
                self.hs["occ"], self.hs["ranges"], fr.sp, self.hs["keys"])
        self.band = oracle.band(self.hs["pos"], self.flags, fr.sc, h)
        self.index = oracle.compact(self.band)
        self.phi, self.wsum = oracle.field_zb05(self.index, self.hs["pos"], self.hs["ranges"],
                                                fr.sp, fr.sc, h * h,
                                                np.float32(fr.particle_radius))


class StageYu13:
    """Oracle results of every stage of one anisotropic (Yu13) frame."""

    def __init__(self, fr, aniso):
        self.fr = fr
        f = np.float32
        h = f(fr.h)
        r = f(h * f(2.0))
        self.hs = hs = oracle.hash_sort(fr.pos, fr.dens, fr.sp)
        self.surf_cell, self.flags, self.involve = oracle.surface_occupancy(
            hs["occ"], hs["ranges"], fr.sp, hs["keys"], involve=True)
        self.band = oracle.band(hs["pos"], self.flags, fr.sc, h)
        self.index = oracle.compact(self.band)
        self.mean, self.smooth = oracle.smooth(hs["pos"], hs["ranges"], fr.sp, r,
                                               f(aniso["lam"]))
        self.inv_idx, self.inv_off = oracle.involved_list(hs["keys"], hs["ranges"], self.involve)
        self.G, self.sigma, lo, hi = oracle.aniso_matrices(
            hs["pos"], self.mean, hs["dens"], self.inv_idx, hs["ranges"], fr.sp, h, r,
            aniso["kr"], aniso["kn"], aniso["ne"])
        ne = aniso["ne"]
        self.ambiguous_inv = (lo < ne) & (hi >= ne)
        amb = np.zeros(fr.n, np.uint8)
        amb[self.inv_idx[self.ambiguous_inv]] = 1
        self.phi, self.flag = oracle.field_aniso(self.index, self.smooth, hs["dens"],
                                                 hs["ranges"], self.involve, self.inv_off,
                                                 self.G, fr.sp, fr.sc, f(fr.mass), amb)


def gpu_buffers(rec, fr, st):
    n, C = fr.n, fr.sp.count
    rx, ry, rz = fr.sc.res
    W = ((rx + 31) // 32) * ry * rz
    B = st["num_band_vertices"]
    return dict(
        keys=rec.buffer("cell_keys", np.uint32, n),
        perm=rec.buffer("permutation", np.uint32, n),
        ranges=rec.buffer("cell_ranges", np.uint32, 2 * C).reshape(C, 2),
        occ=rec.buffer("occupancy", np.float32, C),
        flags=rec.buffer("surface_flags", np.uint32, n),
        band=unpack_bits(rec.buffer("band_bits", np.uint32, W), fr.sc.res),
        index=rec.buffer("band_index", np.uint32, B),
        nverts=rec.buffer("cube_nverts", np.uint32, B),
    )


def field_ok(phi_gpu, phi_ref, h, wsum=None, index=None):
    """Returns (number of violations, number of exempt vertices).  Band
    vertices whose FP64 weight sum is below 1e-6 (support at the kernel edge,
    where sum w > 0 may be decided differently) are exempt."""
    tol = FIELD_RTOL * np.maximum(np.abs(phi_ref), h)
    bad = np.abs(phi_gpu.astype(np.float64) - phi_ref) > tol
    exempt = 0
    if wsum is not None:
        weak = index[wsum < 1e-6]
        exempt = int(bad[weak].sum())
        bad[weak] = False
    return int(bad.sum()), exempt


def compare_mesh(fr, pos_gpu, nrm_gpu, pos_ref, nrm_ref, nrm_len=None):
    cell = float(fr.sc.cell)
    pg = pos_gpu.astype(np.float64)
    mag = np.maximum(np.abs(pg), np.abs(pos_ref))
    ulp = np.spacing(mag.astype(np.float32)).astype(np.float64)
    dv = np.abs(pg - pos_ref) / cell
    vbad = dv > VERTEX_TOL + 2.0 * ulp / cell
    ng = nrm_gpu.astype(np.float64)
    both_nan = np.isnan(ng) & np.isnan(nrm_ref)
    dn = np.where(both_nan, 0.0, np.abs(ng - nrm_ref))
    tol = NORMAL_TOL
    if nrm_len is not None:
        tol = NORMAL_TOL + 2e-6 / np.maximum(nrm_len, 1e-30)[:, None]
    nbad = ~(dn <= tol)
    return int(vbad.any(axis=1).sum()), int(nbad.any(axis=1).sum()), float(np.nanmax(dv, initial=0))
