"""Seeded synthetic particle frames and grid set-up.

The reference reads SPH frames from ``.xyz`` files (GPU/ReconstructorGPU.cpp:200-290)
whose header carries the AABB, kernel radius, particle radius and mass.  The
workloads here are generated instead, deterministically:

* random numbers: splitmix64 (counter based: output k of seed s is
  mix(s + (k + 1) * 0x9E3779B97F4A7C15)); uniform = (u64 >> 40) * 2^-24;
  Gaussian = Box-Muller in FP64 from two uniforms, cast to FP32 at the end;
* physical constants: density 1000, mass = 1000 * s^3, particle radius 0.5 s,
  kernel radius h = 4 s, for lattice spacing s;
* grids: the spatial hashing grid follows the reference's rule
  (GPU/ReconstructorGPU.cpp:237-267: cell = h * scale, AABB padded by one cell,
  res = uint((max - min) / cell + 1), all FP32); the scalar grid is either the
  reference's rule (cell = spatial cell / ratio, same min, same resolution rule)
  or an explicit cubic grid over the padded AABB (cell = max extent / (res - 1)).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
BASE_SEED = 20261201


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


class SplitMix64:
    """Counter-based splitmix64 stream."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def u64(self, n: int) -> np.ndarray:
        k = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            return _mix(self.seed + k * GOLDEN)

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1) with 24 random bits each."""
        return (self.u64(n) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)

    def normal(self, n: int) -> np.ndarray:
        u = self.uniform(2 * n).reshape(n, 2)
        return np.sqrt(-2.0 * np.log(1.0 - u[:, 0])) * np.cos(2.0 * math.pi * u[:, 1])


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Grid:
    min: np.ndarray   # float32[3]
    max: np.ndarray   # float32[3]
    cell: np.float32
    res: tuple

    @property
    def count(self) -> int:
        return int(self.res[0]) * int(self.res[1]) * int(self.res[2])

    def as_tuple(self):
        return (float(self.min[0]), float(self.min[1]), float(self.min[2]), float(self.cell),
                int(self.res[0]), int(self.res[1]), int(self.res[2]))


def spatial_grid(aabb_min, aabb_max, h: float, scale: float = 1.0) -> Grid:
    """GPU/ReconstructorGPU.cpp:239, :254-267 in FP32."""
    f = np.float32
    cell = f(f(h) * f(scale))
    mn = np.asarray(aabb_min, dtype=np.float32) - cell
    mx = np.asarray(aabb_max, dtype=np.float32) + cell
    q = (mx - mn) / cell + f(1.0)
    res = tuple(int(np.uint32(v)) for v in q.astype(np.float32))
    return Grid(mn.astype(np.float32), mx.astype(np.float32), cell, res)


def scalar_grid_ratio(sp: Grid, ratio: int) -> Grid:
    """GPU/ReconstructorGPU.cpp:99-105 in FP32."""
    f = np.float32
    cell = f(sp.cell / f(ratio))
    q = (sp.max - sp.min) / cell + f(1.0)
    res = tuple(int(np.uint32(v)) for v in q.astype(np.float32))
    return Grid(sp.min.copy(), sp.max.copy(), cell, res)


def scalar_grid_cubic(sp: Grid, res: int, cell: float | None = None) -> Grid:
    """Cubic res^3 grid at the padded spatial min; cell = max extent / (res - 1)
    unless given."""
    f = np.float32
    if cell is None:
        ext = float(np.max(sp.max - sp.min))
        cell = ext / (res - 1)
    c = f(cell)
    mx = sp.min + c * f(res - 1)
    return Grid(sp.min.copy(), mx.astype(np.float32), c, (res, res, res))


# ---------------------------------------------------------------------------
# frames
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Frame:
    name: str
    pos: np.ndarray        # float32 (N, 3)
    dens: np.ndarray       # float32 (N,)
    spacing: float
    h: float
    particle_radius: float
    mass: float
    sp: Grid
    sc: Grid
    iso: float
    method: str = "zb05"
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def physical(spacing: float):
    f = np.float32
    s = f(spacing)
    return dict(h=float(f(4.0) * s), particle_radius=float(f(0.5) * s),
                mass=float(f(1000.0) * s * s * s), density=1000.0)


def _finish(name, pos, spacing, sc_mode, iso, method, seed, scale=1.0, meta=None) -> Frame:
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    ph = physical(spacing)
    dens = np.full(pos.shape[0], ph["density"], dtype=np.float32)
    aabb_min = pos.min(axis=0) if len(pos) else np.zeros(3, np.float32)
    aabb_max = pos.max(axis=0) if len(pos) else np.ones(3, np.float32)
    sp = spatial_grid(aabb_min, aabb_max, ph["h"], scale)
    kind, val = sc_mode
    if kind == "ratio":
        sc = scalar_grid_ratio(sp, val)
    elif kind == "cubic":
        sc = scalar_grid_cubic(sp, val)
    elif kind == "cubic_cell":
        sc = scalar_grid_cubic(sp, val[0], val[1])
    else:
        raise ValueError(kind)
    m = {"seed": seed}
    if meta:
        m.update(meta)
    return Frame(name, pos, dens, spacing, ph["h"], ph["particle_radius"], ph["mass"], sp, sc,
                 iso, method, m)


def lattice(nx: int, ny: int, nz: int, spacing: float, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Lattice points origin + (i, j, k) * spacing, x fastest, as float64."""
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    zz, yy, xx = np.meshgrid(k, j, i, indexing="ij")
    p = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1) * spacing
    return p + np.asarray(origin, dtype=np.float64)


def jittered_lattice(n_side: int = 40, spacing: float = 0.01, jitter: float = 0.1,
                     seed: int = BASE_SEED + 0, grid_res: int = 128, iso: float = -1e-4,
                     method: str = "zb05", scale: float = 1.0, cell=None) -> Frame:
    """Plain method, small: n_side^3 lattice at `spacing`, each coordinate
    jittered by U(-jitter, +jitter) * spacing; h = 4 s, hash cell = h * scale;
    scalar grid grid_res^3 with cell 0.5 s (or `cell`) at the padded spatial min."""
    rng = SplitMix64(seed)
    p = lattice(n_side, n_side, n_side, spacing)
    u = rng.uniform(3 * p.shape[0]).reshape(-1, 3)
    p = p + (2.0 * u - 1.0) * (jitter * spacing)
    c = 0.5 * spacing if cell is None else cell
    return _finish("lattice%d" % n_side, p, spacing, ("cubic_cell", (grid_res, c)), iso, method,
                   seed, scale)


def yu13_params(fr: Frame, config_mode: bool = True) -> dict:
    """Anisotropic constants: the configuration values (lambda 0.9, kn 0.5,
    kr 4, Ne 25) or the reference's (lambda 0.5, kn 1, kr 4, Ne 25,
    GPU/ReconstructorGPUOursYu13.cpp:78-99 and ReconstructionCUDA.cu:772,796)."""
    if config_mode:
        return dict(lam=0.9, kn=0.5, kr=4.0, ks=1400.0, ne=25)
    return dict(lam=0.5, kn=1.0, kr=4.0, ks=1.0, ne=25)


def sphere_rejection(rng: SplitMix64, center, radius: float, count: int) -> np.ndarray:
    """`count` points uniform in a ball by rejection from its bounding cube;
    candidates are drawn 3 uniforms at a time and accepted in order."""
    out = []
    have = 0
    c = np.asarray(center, dtype=np.float64)
    while have < count:
        m = max(1024, int((count - have) * 2.2))
        u = rng.uniform(3 * m).reshape(m, 3)
        q = c + (2.0 * u - 1.0) * radius
        ok = np.sum((q - c) ** 2, axis=1) <= radius * radius
        q = q[ok]
        out.append(q)
        have += q.shape[0]
    return np.concatenate(out, axis=0)[:count]


def dam_break(n_target: int = 1_000_000, spacing: float = 0.005, seed: int = BASE_SEED + 1,
              grid_res: int = 256, method: str = "zb05", scale: float = 1.0,
              iso: float = -1e-4, block=(160, 80, 80), sphere_center=(0.4, 0.7, 0.2),
              sphere_radius: float = 0.1, noise_prob: float = 0.1,
              noise_sigma: float = 0.2) -> Frame:
    """Dam-break shape: a lattice block (x fastest) plus a ball of the same
    density (rejection sampled) whose lowest point is 0.2 above the block's top
    face; each particle gets N(0, (noise_sigma*s)^2) noise per coordinate with
    probability noise_prob; then a seeded uniform subset of exactly n_target
    particles, kept in generation order."""
    rng = SplitMix64(seed)
    blk = lattice(block[0], block[1], block[2], spacing)
    n_sph = int(round(4.0 / 3.0 * math.pi * sphere_radius ** 3 / spacing ** 3))
    sph = sphere_rejection(rng, sphere_center, sphere_radius, n_sph)
    p = np.concatenate([blk, sph], axis=0)
    n_all = p.shape[0]
    coin = rng.uniform(n_all)
    z = rng.normal(3 * n_all).reshape(n_all, 3)
    p = p + np.where((coin < noise_prob)[:, None], noise_sigma * spacing * z, 0.0)
    if n_target is not None and n_target < n_all:
        keys = rng.u64(n_all)
        idx = np.sort(np.argsort(keys, kind="stable")[:n_target])
        p = p[idx]
    return _finish("dambreak%d" % p.shape[0], p, spacing, ("cubic", grid_res), iso, method, seed,
                   scale, meta={"generated": n_all, "sphere": n_sph})


def lattice_ball(n_target: int = 256_000, radius: float = 0.25, seed: int = BASE_SEED + 4,
                 grid_res: int = 192, jitter: float = 0.1, iso: float = -1e-4) -> Frame:
    """Jittered lattice ball holding exactly n_target particles: the spacing
    s = radius * (4 pi / 3 / n)^(1/3), lattice points closest to the centre."""
    s = radius * (4.0 * math.pi / 3.0 / n_target) ** (1.0 / 3.0)
    m = int(math.ceil(radius / s)) + 2
    p = lattice(2 * m + 1, 2 * m + 1, 2 * m + 1, s, origin=(-m * s, -m * s, -m * s))
    d = np.sum(p * p, axis=1)
    idx = np.sort(np.argsort(d, kind="stable")[:n_target])
    p = p[idx]
    rng = SplitMix64(seed)
    u = rng.uniform(3 * p.shape[0]).reshape(-1, 3)
    p = p + (2.0 * u - 1.0) * (jitter * s)
    return _finish("ball%d" % n_target, p, s, ("cubic", grid_res), iso, "zb05", seed)


def random_shapes_frame(index: int, seed: int = BASE_SEED + 3, n_min: float = 2e4,
                        n_max: float = 2e6, res_min: int = 96, res_max: int = 320,
                        method: str = "zb05", iso: float = -1e-4, scale: float = 1.0,
                        n_override: int | None = None) -> Frame:
    """One frame of the batched workload: stream seed + index; a particle count
    N log-uniform in [n_min, n_max]; 1-4 shapes (block or ball, 50:50) at random
    sizes and positions in the unit cube; spacing s = (sum of shape volumes /
    N)^(1/3); every shape filled with the lattice of spacing s anchored at its
    minimum corner (balls keep the lattice points inside), jittered by
    U(+-0.1 s); a cubic scalar grid at 0.5 s over the padded AABB, its
    resolution clamped to [res_min, res_max]."""
    rng = SplitMix64(seed + index)
    u = rng.uniform(1)[0]
    n_target = n_override or int(round(math.exp(math.log(n_min) + u * (math.log(n_max) - math.log(n_min)))))
    nshape = 1 + int(rng.uniform(1)[0] * 4)
    shapes = []
    vol = 0.0
    for _ in range(nshape):
        kind, a, b, c, d, e, f = rng.uniform(7)
        if kind < 0.5:
            size = 0.15 + 0.45 * np.array([a, b, c])
            lo = np.array([d, e, f]) * (1.0 - size)
            shapes.append(("block", lo, size))
            vol += float(np.prod(size))
        else:
            r = 0.08 + 0.22 * a
            cen = r + np.array([b, c, d]) * (1.0 - 2 * r)
            shapes.append(("ball", cen, r))
            vol += 4.0 / 3.0 * math.pi * r ** 3
    s = (vol / n_target) ** (1.0 / 3.0)
    pts = []
    for kind, p, q in shapes:
        if kind == "block":
            n = np.maximum(1, np.floor(q / s).astype(int))
            pts.append(lattice(n[0], n[1], n[2], s, origin=p))
        else:
            m = int(math.ceil(2 * q / s)) + 1
            g = lattice(m, m, m, s, origin=p - q)
            pts.append(g[np.sum((g - p) ** 2, axis=1) <= q * q])
    pos = np.concatenate(pts, axis=0)
    pos = pos + (2.0 * rng.uniform(3 * pos.shape[0]).reshape(-1, 3) - 1.0) * (0.1 * s)
    h = 4.0 * s
    sp = spatial_grid(pos.min(0).astype(np.float32), pos.max(0).astype(np.float32),
                      float(np.float32(h)), scale)
    ext = float(np.max(sp.max - sp.min))
    res = int(min(max(math.ceil(ext / (0.5 * s)) + 1, res_min), res_max))
    fr = _finish("shapes%02d" % index, pos, s, ("cubic", res), iso, method, seed + index, scale,
                 meta={"target": n_target, "shapes": [k for k, _, _ in shapes]})
    return fr


def small_block(nx=12, ny=10, nz=8, spacing=0.01, seed=BASE_SEED + 100, ratio=2, jitter=0.1,
                iso=-1e-4, method="zb05") -> Frame:
    """Small jittered block on the reference's own grid rule (scalar cell =
    hash cell / ratio), for parity tests of the reference-rule grids."""
    rng = SplitMix64(seed)
    p = lattice(nx, ny, nz, spacing)
    u = rng.uniform(3 * p.shape[0]).reshape(-1, 3)
    p = p + (2.0 * u - 1.0) * (jitter * spacing)
    return _finish("block%dx%dx%d" % (nx, ny, nz), p, spacing, ("ratio", ratio), iso, method, seed)


def scattered(n: int = 3000, extent: float = 1.6, spacing: float = 0.01,
              seed: int = BASE_SEED + 101, ratio: int = 1, iso: float = -1e-4,
              method: str = "zb05") -> Frame:
    """Sparse particles, uniform in a cube of side `extent`, on the reference's
    grid rule with scalar cell = hash cell / ratio: few particles per hash cell,
    so 64 consecutive sorted particles span many cells (coarse-grid band path
    with widely spread exceptions)."""
    rng = SplitMix64(seed)
    p = rng.uniform(3 * n).reshape(-1, 3) * extent
    return _finish("scattered%d" % n, p, spacing, ("ratio", ratio), iso, method, seed)


def sim_param_fields(fr: Frame, ratio: int = 2, lam: float = 0.5, scale: float = 1.0,
                     min_neighbors: int = 25) -> dict:
    f = np.float32
    h = f(fr.h)
    return dict(
        particle_radius=float(f(fr.particle_radius)), particle_mass=float(f(fr.mass)),
        smoothing_radius=float(h), smoothing_radius_sq=float(h * h),
        smoothing_radius_inv=float(f(1.0 / float(h))), lambda_for_smoothed=float(f(lam)),
        anisotropic_radius=float(h * f(2.0)), iso_value=float(f(fr.iso)),
        sc_sp_grid_res_ratio=int(ratio), expand_extent=4, min_num_neighbors=int(min_neighbors),
        spatial_cell_size_scale=float(f(scale)))
