"""Benchmark: full reconstruction frames/s on one MI355X.

Default workload: dam-break frame of 1,000,000 particles (a 160x80x80 lattice
block at spacing 0.005 plus a ball of radius 0.1 at the same density, 10% of
the particles with Gaussian noise, seeded subset), cubic 256^3 scalar grid over
the padded AABB, ZB05 field on the occupancy band (the reference's
ReconstructorGPUOursZB05 stage order).  A step = one whole frame: spatial
hashing -> surface particles -> band -> field -> marching cubes, inputs
resident in HBM (each step reads its own unsorted copy, since a frame sorts its
particles in place).

Prints one JSON line (see the `roofline` and `cpu_baseline` objects).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("frames/s at 1M particles, 256^3 grid; field-stage HBM GB/s; batched vs per-frame")
HBM_PEAK = 8.0e12
FP32_PEAK = 157.3e12


def field_counts(fr, band_index: np.ndarray, ranges: np.ndarray):
    """Algorithmic quantities of the ZB05 field stage, from the band and the
    cell ranges: P27 = sum over band vertices of the particles in the clamped
    3^3 cell block of the vertex's hash cell; N_touched = particles in the union
    of those blocks."""
    rx, ry, rz = fr.sp.res
    cnt = np.where(ranges[:, 0] == 0xFFFFFFFF, 0, ranges[:, 1].astype(np.int64) - ranges[:, 0])
    c3 = cnt.reshape(rz, ry, rx)
    sat = np.zeros((rz + 1, ry + 1, rx + 1), np.int64)
    sat[1:, 1:, 1:] = c3.cumsum(0).cumsum(1).cumsum(2)
    sx, sy, _ = fr.sc.res
    v = band_index.astype(np.int64)
    idx = np.stack([v % sx, (v // sx) % sy, v // (sx * sy)], 1)
    cell = []
    for a, (mn, r) in enumerate(zip((fr.sp.min[0], fr.sp.min[1], fr.sp.min[2]), (rx, ry, rz))):
        vf = (idx[:, a].astype(np.float64) * float(fr.sc.cell) + float(fr.sc.min[a])).astype(np.float32)
        q = (vf - np.float32(mn)) / np.float32(fr.sp.cell)
        c = np.trunc(np.clip(q, -4, r + 4)).astype(np.int64)
        cell.append(c)
    lo = [np.clip(c - 1, 0, r - 1) for c, r in zip(cell, (rx, ry, rz))]
    hi = [np.clip(c + 1, 0, r - 1) + 1 for c, r in zip(cell, (rx, ry, rz))]
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    box = (sat[z1, y1, x1] - sat[z0, y1, x1] - sat[z1, y0, x1] - sat[z1, y1, x0]
           + sat[z0, y0, x1] + sat[z0, y1, x0] + sat[z1, y0, x0] - sat[z0, y0, x0])
    p27 = int(box.sum())
    touched = np.zeros((rz, ry, rx), bool)
    uc = np.unique(np.stack([np.clip(c, 0, r - 1) for c, r in zip(cell, (rx, ry, rz))], 1), axis=0)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                touched[np.clip(uc[:, 2] + dz, 0, rz - 1), np.clip(uc[:, 1] + dy, 0, ry - 1),
                        np.clip(uc[:, 0] + dx, 0, rx - 1)] = True
    n_touched = int(c3[touched].sum())
    return p27, n_touched


def cpu_baseline(fr, method: str, budget_s: float = 25.0):
    """The CPU oracle (FP64, OpenMP on every host core) on the same frame."""
    import oracle

    threads = oracle.threads()
    times = []
    t_total = 0.0
    while len(times) < 3 and t_total < budget_s:
        t0 = time.perf_counter()
        oracle.frame(method, fr)
        dt = time.perf_counter() - t0
        times.append(dt)
        t_total += dt
        if dt > 10.0:
            break
    med = float(np.median(times))
    return {"value": 1.0 / med, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} full frame(s) of the same workload through the FP64 "
                      f"oracle, median {med:.2f} s",
            "nproc": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--method", default="zb05", choices=["zb05", "akinci12", "yu13"])
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=None,
                    help="scalar grid resolution (default 256, 384 for yu13)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus %d is not supported "
                                   "(frames shard across GPUs without communication, see "
                                   "DESIGN.md)" % args.gpus}))
        sys.exit(2)

    import torch

    from fluidsurfacereconstructiongpu_amd import datasets as ds
    from fluidsurfacereconstructiongpu_amd.reconstructor import Reconstructor, stats_dict

    yu13 = args.method == "yu13"
    if args.grid is None:
        args.grid = 384 if yu13 else 256
    t0 = time.perf_counter()

    def make_frame():
        if yu13:
            return ds.dam_break(n_target=args.particles, grid_res=args.grid, method="yu13",
                                scale=1.3, iso=0.0)
        return ds.dam_break(n_target=args.particles, grid_res=args.grid, method=args.method)

    fr = make_frame()
    print(f"[bench] workload generated in {time.perf_counter() - t0:.1f} s: N={fr.n} "
          f"sp={fr.sp.res} sc={fr.sc.res}", file=sys.stderr, flush=True)
    dev = torch.device("cuda", 0)
    rec = Reconstructor(args.method, aniso=ds.yu13_params(fr, True) if yu13 else None)
    K, W = args.steps, args.warmup
    src_pos = torch.from_numpy(fr.pos).to(dev)
    src_dens = torch.from_numpy(fr.dens).to(dev)
    # one resident unsorted copy per step
    pos_k = [src_pos.clone() for _ in range(K + W)]
    dens_k = [src_dens.clone() for _ in range(K + W)]
    first = rec.run(fr, pos=src_pos.clone(), dens=src_dens.clone())
    cap = max(first.stats["num_mesh_vertices"] + 1024, 3 * 1024)
    sim = rec.params(fr)
    stream = torch.cuda.current_stream(dev)
    for i in range(W):
        rec.launch(fr, pos_k[i], dens_k[i], cap, sim=sim, stream=stream)
    torch.cuda.synchronize(dev)
    rec.ctx.set_profiling(K)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize(dev)
    ev0.record(stream)
    for i in range(W, W + K):
        rec.launch(fr, pos_k[i], dens_k[i], cap, sim=sim, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize(dev)
    total_ms = ev0.elapsed_time(ev1)
    stages = rec.ctx.stage_times(K)
    rec.ctx.set_profiling(0)
    st = stats_dict(rec._stats)
    print(f"[bench] {K} frames in {total_ms:.2f} ms; stats {st}", file=sys.stderr, flush=True)

    frame_ms = stages.sum(axis=1)
    field_ms = float(stages[:, 4].mean())
    fill_ms = float(stages[:, 3].mean())
    B = st["num_band_vertices"]
    C, V = fr.sp.count, fr.sc.count
    if yu13:
        roofline = None
    else:
        # algorithmic counts of the field stage
        band_index = rec.buffer("band_index", np.uint32, B)
        ranges = rec.buffer("cell_ranges", np.uint32, 2 * fr.sp.count).reshape(-1, 2)
        p27, n_touched = field_counts(fr, band_index, ranges)
        flops = 20.0 * p27
        b_comp = 8.0 * B + 8.0 * C + 12.0 * n_touched + 4.0 * V
        achieved_tflops = flops / (field_ms * 1e-3) / 1e12
        t_roof = max(b_comp / HBM_PEAK, flops / FP32_PEAK)
        roofline = {"bound": "valu", "kernel": "k_field_zb05", "achieved": achieved_tflops,
                    "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                    "frac": achieved_tflops / (FP32_PEAK / 1e12), "traffic": None,
                    "flops_per_launch": flops, "pairs_p27": p27,
                    "field_stage_roofline_frac": t_roof / ((field_ms + fill_ms) * 1e-3),
                    "field_stage_hbm_gbps": b_comp / ((field_ms + fill_ms) * 1e-3) / 1e9,
                    "compulsory_bytes": b_comp, "field_kernel_ms": field_ms,
                    "fill_kernel_ms": fill_ms}
    out = {
        "metric": METRIC,
        "value": K / (total_ms * 1e-3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": total_ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "dam-break frame, %d particles, %d^3 scalar grid, %s field on the "
                               "occupancy band, full marching cubes" % (fr.n, args.grid,
                                                                         args.method.upper()),
                   "particles": fr.n, "grid": list(fr.sc.res), "hash_grid": list(fr.sp.res),
                   "method": args.method, "parallelism": "single GPU"},
        "roofline": roofline,
        "frame_ms_p50": float(np.percentile(frame_ms, 50)),
        "frame_ms_p99": float(np.percentile(frame_ms, 99)),
        "stage_ms": {k: float(stages[:, i].mean()) for i, k in enumerate(
            ["hash", "surface", "band", "fill", "field", "mc"])},
        "counts": {k: st[k] for k in ("num_surface_particles", "num_band_vertices",
                                      "num_valid_cubes", "num_mesh_vertices")},
    }
    if not args.no_cpu_baseline and not yu13:
        out["cpu_baseline"] = cpu_baseline(make_frame(), args.method)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
