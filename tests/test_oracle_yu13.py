"""Pins of the anisotropic (Yu13 / TC01) part of the CPU oracle by closed forms,
symmetry and a library eigen-solver."""
import math

import numpy as np

import oracle
from fluidsurfacereconstructiongpu_amd import datasets as ds


def grid(minv, cell, res):
    f = np.float32
    mn = np.asarray(minv, np.float32)
    return ds.Grid(mn, (mn + f(cell) * np.asarray(res, np.float32)).astype(np.float32), f(cell),
                   tuple(int(r) for r in res))


def test_jacobi_matches_numpy_eigh():
    rng = np.random.default_rng(5)
    for _ in range(200):
        x = rng.normal(size=(3, 3)) * rng.uniform(1e-3, 1.0, size=3)
        a = x @ x.T + 1e-6 * np.eye(3)
        d, v = oracle.eigen3(a)
        assert np.allclose(np.sort(d), np.linalg.eigvalsh(a), rtol=1e-12, atol=1e-15)
        assert np.allclose(v @ np.diag(d) @ v.T, a, rtol=1e-12, atol=1e-15)
        assert np.allclose(v.T @ v, np.eye(3), atol=1e-12)


def lattice_frame(n=11, s=0.01, scale=1.3):
    p = ds.lattice(n, n, n, s).astype(np.float32)
    h = np.float32(4 * s)
    sp = ds.spatial_grid(p.min(0), p.max(0), h, scale)
    r = oracle.hash_sort(p, np.full(len(p), 1000.0, np.float32), sp)
    return p, h, sp, r


def test_smoothing_lattice_interior_is_fixed_point():
    p, h, sp, r = lattice_frame()
    mean, sm = oracle.smooth(r["pos"], r["ranges"], sp, 2 * h, 0.5)
    centre = np.all(np.abs(r["pos"] - 0.05) < 1e-6, axis=1)
    assert centre.sum() == 1
    i = np.nonzero(centre)[0][0]
    assert np.allclose(mean[i], r["pos"][i], atol=1e-12)
    assert np.allclose(sm[i], r["pos"][i], atol=1e-12)
    # a corner particle is pulled towards the block
    corner = np.nonzero(np.all(np.abs(r["pos"]) < 1e-6, axis=1))[0][0]
    assert np.all(mean[corner] > 0.005)
    assert np.allclose(sm[corner], 0.5 * mean[corner], atol=1e-12)


def test_covariance_isotropic_interior_and_isolated():
    p, h, sp, r = lattice_frame()
    mean, sm = oracle.smooth(r["pos"], r["ranges"], sp, 2 * h, 0.5)
    # centre particle: isotropic covariance -> G = I / h
    i = np.nonzero(np.all(np.abs(r["pos"] - 0.05) < 1e-6, axis=1))[0][0]
    G, sig, lo, hi = oracle.aniso_matrices(r["pos"], mean, r["dens"], np.array([i], np.uint32),
                                           r["ranges"], sp, h, 2 * h)
    assert lo[0] >= 25
    assert np.allclose(G[0], np.eye(3) / float(h), rtol=1e-9)
    assert abs(sig[0, 0] - sig[0, 2]) < 1e-9 * sig[0, 0]
    # covariance by brute force: eigenvalues of (h^2 I + sum w v v^T) / sum w
    P = r["pos"].astype(np.float64)
    v = P - mean[i]
    d = np.linalg.norm(v, axis=1)
    w = np.where(d < 2 * float(h), 1 - (d / (2 * float(h))) ** 3, 0.0)
    C = (float(np.float32(h * h)) * np.eye(3) + (w[:, None, None] * v[:, :, None] * v[:, None, :]).sum(0)) / w.sum()
    assert np.allclose(np.sort(np.linalg.eigvalsh(C))[::-1], sig[0], rtol=1e-12)
    # isolated particle (one neighbour only: itself) -> G = I / (kn h)
    q = np.array([[0.0, 0.0, 0.0]], np.float32)
    sp1 = ds.spatial_grid(q.min(0), q.max(0), h, 1.3)
    r1 = oracle.hash_sort(q, np.ones(1, np.float32), sp1)
    m1, _ = oracle.smooth(r1["pos"], r1["ranges"], sp1, 2 * h, 0.5)
    G1, _, _, _ = oracle.aniso_matrices(r1["pos"], m1, r1["dens"], np.array([0], np.uint32),
                                        r1["ranges"], sp1, h, 2 * h, kn=0.5)
    assert np.allclose(G1[0], np.eye(3) / (0.5 * float(h)))


def test_covariance_flat_face_clamps_normal_direction():
    p, h, sp, r = lattice_frame(n=13)
    mean, _ = oracle.smooth(r["pos"], r["ranges"], sp, 2 * h, 0.5)
    # middle of the z = 0 face
    i = np.nonzero(np.all(np.abs(r["pos"] - [0.06, 0.06, 0.0]) < 1e-6, axis=1))[0][0]
    G, sig, lo, _ = oracle.aniso_matrices(r["pos"], mean, r["dens"], np.array([i], np.uint32),
                                          r["ranges"], sp, h, 2 * h, kr=4.0)
    assert lo[0] >= 25
    ev, evec = np.linalg.eigh(G[0])
    # the stretched direction of G (largest eigenvalue <-> smallest sigma) is the normal z
    assert abs(abs(evec[2, 2]) - 1.0) < 1e-6
    assert sig[0, 2] < sig[0, 0]
    # det(G) = 1 / h^3 (volume preservation) whatever the clamping
    assert abs(np.linalg.det(G[0]) * float(h) ** 3 - 1.0) < 1e-9


def test_aniso_field_isolated_particle_closed_form_and_normalisation():
    h = np.float32(0.04)
    q = np.array([[0.0213, 0.0187, 0.0241]], np.float32)
    sp = grid((-0.1, -0.1, -0.1), 0.052, (5, 5, 5))
    sc = grid((-0.06, -0.06, -0.06), 0.004, (40, 40, 40))
    r = oracle.hash_sort(q, np.full(1, 1000.0, np.float32), sp)
    involve = np.zeros(sp.count, np.uint32)
    c = r["keys"][0]
    involve[c] = 1
    off = np.zeros(sp.count, np.uint32)
    off[c + 1:] = 1
    G = (np.eye(3) / float(h)).reshape(1, 3, 3)
    idx = np.arange(sc.count, dtype=np.uint32)
    mass = np.float32(1000.0 * 0.01 ** 3)
    phi, _ = oracle.field_aniso(idx, r["pos"].astype(np.float64), r["dens"], r["ranges"],
                                involve, off, G, sp, sc, mass)
    x = np.arange(40, dtype=np.float32) * np.float32(0.004) + np.float32(-0.06)
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    d2 = ((xx.astype(np.float64) - q[0, 0]) ** 2 + (yy - q[0, 1]) ** 2 + (zz - q[0, 2]) ** 2) / float(h) ** 2
    sigma = 315 / (64 * math.pi)
    W = np.where(d2 < 1, sigma / float(h) ** 3 * (1 - d2) ** 3, 0.0)
    mr = float(mass) / 1000.0
    # the vertices whose hash cell block holds the particle see the closed form
    assert np.allclose(phi, (mr * W - 0.5).ravel(), atol=1e-12)
    # integral of W is 1: sum (phi + 0.5) dV = m / rho
    assert abs((phi + 0.5).sum() * 0.004 ** 3 / mr - 1.0) < 2e-3


def test_aniso_field_non_involved_uses_unscaled_kernel():
    h = np.float32(0.04)
    q = np.array([[0.0213, 0.0187, 0.0241]], np.float32)
    sp = grid((-0.1, -0.1, -0.1), 0.052, (5, 5, 5))
    sc = grid((-0.06, -0.06, -0.06), 0.004, (40, 40, 40))
    r = oracle.hash_sort(q, np.full(1, 1000.0, np.float32), sp)
    involve = np.zeros(sp.count, np.uint32)
    off = np.zeros(sp.count, np.uint32)
    idx = np.array([0, 12345, sc.count - 1], np.uint32)
    mass = np.float32(1e-3)
    phi, _ = oracle.field_aniso(idx, r["pos"].astype(np.float64), r["dens"], r["ranges"],
                                involve, off, np.zeros((0, 3, 3)), sp, sc, mass)
    sigma = 315 / (64 * math.pi)
    for v in idx:
        x, y, z = v % 40, (v // 40) % 40, v // 1600
        vp = np.array([np.float32(np.float32(t) * np.float32(0.004) + np.float32(-0.06))
                       for t in (x, y, z)], np.float64)
        rr = ((q[0].astype(np.float64) - vp) ** 2).sum()
        # world units, no 1/h, no det: the particle is seen from every vertex whose
        # 5^3 block contains its cell
        cell_v = np.trunc((vp.astype(np.float32) - sp.min) / sp.cell).astype(int)
        cell_p = np.trunc((q[0] - sp.min) / sp.cell).astype(int)
        seen = np.all(np.abs(cell_v - cell_p) <= 2)
        exp = (float(mass) / 1000.0 * sigma * (1 - rr) ** 3 if seen else 0.0) - 0.5
        assert abs(phi[v] - exp) < 1e-12
