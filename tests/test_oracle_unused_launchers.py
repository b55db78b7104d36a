"""Oracle pins of the two launchers the reference defines but never calls:
estimateSurfaceVertices (ReconstructionCUDA.cu:123-175) and the isotropic TC01
field (ReconstructionCUDA.cu:178-227, kernelTC01 ReconstructionCUDA.cuh:252-260).
"""
import numpy as np

import oracle
from fluidsurfacereconstructiongpu_amd import datasets as ds


def grid(minv, cell, res):
    f = np.float32
    mn = np.asarray(minv, np.float32)
    return ds.Grid(mn, (mn + f(cell) * np.asarray(res, np.float32)).astype(np.float32), f(cell),
                   tuple(int(r) for r in res))


def brute_estimate(dens, dg, ratio, ext, sg):
    rx, ry, rz = dg.res
    d = dens.reshape(rz, ry, rx)
    out = np.zeros((sg.res[2], sg.res[1], sg.res[0]), np.uint32)
    for z in range(rz - 1):
        for y in range(ry - 1):
            for x in range(rx - 1):
                c = d[z:z + 2, y:y + 2, x:x + 2] <= 0.5
                if c.all() or not c.any():
                    continue
                lo = [min(max(v * ratio - ext, 0), r - 1) for v, r in zip((x, y, z), sg.res)]
                hi = [min(max((v + 1) * ratio + ext, 0), r - 1) for v, r in zip((x, y, z), sg.res)]
                out[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = 1
    return out.ravel()


def test_estimate_single_cell_box():
    dg = grid((0, 0, 0), 1.0, (5, 5, 5))
    sg = grid((0, 0, 0), 0.5, (11, 11, 11))
    dens = np.zeros(dg.count, np.float32)
    dens[(2 * 5 + 2) * 5 + 2] = 1.0            # one occupied cell
    out = oracle.estimate_surface_vertices(dens, dg, 2, 1, sg).reshape(11, 11, 11)
    # the 8 cells that have the occupied cell as a corner each mark
    # [c*2 - 1, (c+1)*2 + 1] -> the union spans [1*2-1, (2+1)*2+1] = [1, 7] per axis
    idx = np.argwhere(out)
    assert idx.min(axis=0).tolist() == [1, 1, 1] and idx.max(axis=0).tolist() == [7, 7, 7]
    assert out.sum() == 7 ** 3


def test_estimate_brute_force():
    rng = np.random.default_rng(3)
    dg = grid((0, 0, 0), 1.0, (7, 6, 5))
    sg = grid((0, 0, 0), 1.0 / 3, (19, 16, 13))
    dens = (rng.uniform(size=dg.count) < 0.4).astype(np.float32)
    for ratio, ext in ((3, 3), (2, 0), (3, 5)):
        got = oracle.estimate_surface_vertices(dens, dg, ratio, ext, sg)
        assert np.array_equal(got, brute_estimate(dens, dg, ratio, ext, sg))


def test_tc01_single_particle_closed_form():
    sp = grid((0, 0, 0), 0.1, (5, 5, 5))
    sc = grid((0, 0, 0), 0.025, (21, 21, 21))
    p = np.array([[0.2, 0.21, 0.19]], np.float32)
    hs = oracle.hash_sort(p, np.ones(1, np.float32), sp)
    idx = np.arange(sc.count, dtype=np.uint32)
    R2 = np.float32(0.1) * np.float32(0.1)
    phi = oracle.field_tc01(idx, hs["pos"], hs["ranges"], sp, sc, R2)
    g = (np.arange(21, dtype=np.float32) * np.float32(0.025)).astype(np.float64)
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    d2 = ((xx - float(p[0, 0])) ** 2 + (yy - float(p[0, 1])) ** 2 +
          (zz - float(p[0, 2])) ** 2).ravel()
    ref = np.where(d2 <= float(R2), (d2 / float(R2) * 0.5 - 0.5) ** 2, 0.0)
    np.testing.assert_allclose(phi, ref, rtol=1e-12, atol=1e-15)
    assert abs(phi.max() - 0.25) < 0.01


def test_tc01_brute_force_all_particles():
    fr = ds.jittered_lattice(n_side=8, grid_res=40)
    hs = oracle.hash_sort(fr.pos, fr.dens, fr.sp)
    idx = np.arange(0, fr.sc.count, 7, dtype=np.uint32)
    h = np.float32(fr.h)
    phi = oracle.field_tc01(idx, hs["pos"], hs["ranges"], fr.sp, fr.sc, h * h)
    rx, ry = fr.sc.res[0], fr.sc.res[1]
    v = idx.astype(np.int64)
    pts = np.stack([v % rx, (v // rx) % ry, v // (rx * ry)], 1)
    # vertex positions as one fused multiply-add: exact product and sum in
    # FP64, a single rounding to FP32
    vpos = np.stack([(pts[:, a].astype(np.float64) * float(fr.sc.cell) + float(fr.sc.min[a]))
                     .astype(np.float32) for a in range(3)], 1).astype(np.float64)
    P = hs["pos"].astype(np.float64)
    d2 = ((vpos[:, None, :] - P[None, :, :]) ** 2).sum(-1)
    R2 = float(h * h)
    ref = np.where(d2 <= R2, (d2 / R2 * 0.5 - 0.5) ** 2, 0.0).sum(1)
    # the 3^3 block holds every particle within R because the hash cell is R
    np.testing.assert_allclose(phi[idx], ref, rtol=1e-9, atol=1e-12)
