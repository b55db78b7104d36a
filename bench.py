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


def batched_vs_per_frame(method: str, frames: int, lanes: int, reps: int, autotune: bool):
    """The batched workload: `frames` random-shape frames (particle counts
    log-uniform in [2e4, 2e6], 96^3-320^3 grids), each starting from its own
    resident unsorted copy.  Per-frame path: the frames one after the other on
    one stream (copy + frame each).  Batched path: the same work through
    fsr_batch_run (lanes concurrent, one graph launch per pass)."""
    import torch

    from fluidsurfacereconstructiongpu_amd import datasets as ds
    from fluidsurfacereconstructiongpu_amd.reconstructor import Batch

    kw = dict(method="yu13", iso=0.0, scale=1.3) if method == "yu13" else {}
    t0 = time.perf_counter()
    frs = [ds.random_shapes_frame(i, **kw) for i in range(frames)]
    print(f"[bench] batch workload: {frames} frames, {sum(f.n for f in frs)} particles, "
          f"generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    aniso = ds.yu13_params(frs[0], True) if method == "yu13" else None
    batch = Batch(method, frs, lanes=lanes, aniso=aniso)
    dev = batch.dev
    stream = torch.cuda.current_stream(dev)
    tuned = None
    if autotune:
        import ctypes
        from fluidsurfacereconstructiongpu_amd import _lib

        big = int(np.argmax([f.n for f in frs]))
        choice = (ctypes.c_uint32 * 4)()
        ap = batch.rec.aniso_param()
        _lib.check(batch.ctx.lib.fsr_autotune(batch.ctx.handle, {"zb05": 0, "akinci12": 1, "yu13": 2}[method],
                                              ctypes.byref(ap), ctypes.byref(batch.descs[big]),
                                              5, choice, ctypes.c_void_p(stream.cuda_stream)))
        tuned = list(choice)
    # per-frame path on the batch's parent context: copy + frame, one stream
    lib, ctx = batch.ctx.lib, batch.ctx.handle
    from fluidsurfacereconstructiongpu_amd.reconstructor import METHODS
    ap = batch.rec.aniso_param()
    import ctypes

    def per_frame_pass(events=None, sync=False):
        for i, d in enumerate(batch.descs):
            b = batch.bufs[i]
            if events is not None:
                events[i][0].record(stream)
            b["pos"].copy_(b["src_pos"])
            b["dens"].copy_(b["src_dens"])
            st = lib.fsr_reconstruct_frame(ctx, METHODS[method], ctypes.byref(d.sim),
                                           ctypes.byref(ap), ctypes.byref(d.sp_grid),
                                           ctypes.byref(d.sc_grid), d.pos, d.dens,
                                           d.num_particles, d.phi, d.tri_pos, d.tri_nrm,
                                           d.tri_capacity, d.stats,
                                           ctypes.c_void_p(stream.cuda_stream))
            if st != 0:
                raise RuntimeError("frame %d failed: %d" % (i, st))
            if events is not None:
                events[i][1].record(stream)
            if sync:
                # a per-frame consumer reads the frame's counts before using its mesh
                int(b["stats"][6].item())

    per_frame_pass()                       # sizes the arena
    torch.cuda.synchronize(dev)
    if batch.grow_overflowed():
        per_frame_pass()
        torch.cuda.synchronize(dev)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in frs]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for r in range(reps):
        per_frame_pass(evs if r == reps - 1 else None)
    e1.record(stream)
    torch.cuda.synchronize(dev)
    per_ms = e0.elapsed_time(e1) / reps
    per_frame_ms = np.array([a.elapsed_time(b) for a, b in evs])
    # per-frame path with the count read back after every frame (host sync)
    t_sync = time.perf_counter()
    for r in range(reps):
        per_frame_pass(sync=True)
    torch.cuda.synchronize(dev)
    sync_ms = (time.perf_counter() - t_sync) * 1e3 / reps
    # batched path
    batch.run(stream=stream)                # eager + capture
    torch.cuda.synchronize(dev)
    batch.run(stream=stream)                # first replay
    torch.cuda.synchronize(dev)
    e0.record(stream)
    for _ in range(reps):
        batch.run(stream=stream)
    e1.record(stream)
    torch.cuda.synchronize(dev)
    bat_ms = e0.elapsed_time(e1) / reps
    bat_frame_ms = batch.frame_ms()
    batch.close()
    n = len(frs)
    return {
        "workload": "%d random-shape frames (%s), %d particles in total, grids %d^3-%d^3"
                    % (n, method, sum(f.n for f in frs), min(f.sc.res[0] for f in frs),
                       max(f.sc.res[0] for f in frs)),
        "lanes": lanes, "passes": reps, "autotuned_workgroups_per_cu": tuned,
        "batched_frames_per_s": n / (bat_ms * 1e-3),
        "per_frame_sync_frames_per_s": n / (sync_ms * 1e-3),
        "per_frame_async_frames_per_s": n / (per_ms * 1e-3),
        "speedup_vs_per_frame_sync": sync_ms / bat_ms,
        "speedup_vs_per_frame_async": per_ms / bat_ms,
        "batched_frame_ms_p50": float(np.percentile(bat_frame_ms, 50)),
        "batched_frame_ms_p99": float(np.percentile(bat_frame_ms, 99)),
        "per_frame_ms_p50": float(np.percentile(per_frame_ms, 50)),
        "per_frame_ms_p99": float(np.percentile(per_frame_ms, 99)),
    }


def adjoint_cost(reps: int = 20):
    """Adjoint workload: 256,000-particle jittered lattice ball (radius 0.25),
    192^3 grid, ZB05, sphere loss (r = 0.3) over the grid and the soup.  Device
    time of the forward frame against the backward (vertex term + particle
    gather), both on the same stream, inputs resident."""
    import ctypes

    import torch

    from fluidsurfacereconstructiongpu_amd import _lib
    from fluidsurfacereconstructiongpu_amd import datasets as ds
    from fluidsurfacereconstructiongpu_amd.autograd import SurfaceField, sphere_loss

    fr = ds.lattice_ball()
    sf = SurfaceField(fr, "zb05")
    dev = sf.device
    src = torch.from_numpy(fr.pos).to(dev)
    phi, soup = sf.forward_frame(src)
    ph = phi.detach().clone().requires_grad_(True)
    so = soup.detach().clone().requires_grad_(True)
    sphere_loss(ph, so, fr).backward()
    g_phi0 = ph.grad.to(torch.float32).contiguous()
    g_soup = so.grad.to(torch.float32).contiguous()
    stream = torch.cuda.current_stream(dev)
    s = ctypes.c_void_p(stream.cuda_stream)
    lib, h = sf.ctx.lib, sf.ctx.handle
    pos_k = [src.clone() for _ in range(reps + 1)]
    dens_k = [sf.dens.clone() for _ in range(reps + 1)]
    g_k = [g_phi0.clone() for _ in range(2 * reps + 1)]
    out = torch.empty((fr.n, 3), dtype=torch.float32, device=dev)

    def bwd(g, mc=True):
        if mc:
            _lib.check(lib.fsr_mc_vertex_backward(h, ctypes.c_void_p(g_soup.data_ptr()),
                                                  ctypes.c_void_p(g.data_ptr()), s))
        _lib.check(lib.fsr_field_zb05_backward(h, ctypes.c_void_p(g.data_ptr()),
                                               ctypes.c_void_p(out.data_ptr()), s))

    sf.rec.launch(fr, pos_k[reps], dens_k[reps], sf.cap, sim=sf.sim, stream=stream)
    bwd(g_k[2 * reps])
    torch.cuda.synchronize(dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(stream)
    for i in range(reps):
        sf.rec.launch(fr, pos_k[i], dens_k[i], sf.cap, sim=sf.sim, stream=stream)
    ev[1].record(stream)
    for i in range(reps):
        bwd(g_k[i])
    ev[2].record(stream)
    for i in range(reps):
        bwd(g_k[reps + i], mc=False)
    ev[3].record(stream)
    torch.cuda.synchronize(dev)
    t_fwd = ev[0].elapsed_time(ev[1]) / reps
    t_bwd = ev[1].elapsed_time(ev[2]) / reps
    t_field = ev[2].elapsed_time(ev[3]) / reps
    st = stats_dict_of(sf)
    res = {"workload": "jittered lattice ball, %d particles, 192^3 grid, ZB05, sphere loss "
                       "over grid and soup" % fr.n,
           "t_fwd_ms": t_fwd, "t_bwd_ms": t_bwd, "t_bwd_field_ms": t_field,
           "t_bwd_vertex_term_ms": t_bwd - t_field, "ratio_bwd_fwd": t_bwd / t_fwd,
           "band_vertices": st["num_band_vertices"], "soup_vertices": st["num_mesh_vertices"],
           "reps": reps}
    sf.close()
    return res


def stats_dict_of(sf):
    from fluidsurfacereconstructiongpu_amd.reconstructor import stats_dict

    return stats_dict(sf.rec._stats)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_field_zb05_dambreak1M_256_hbm_traffic.md")


def measured_traffic(args):
    """HBM bytes per k_field_zb05 launch on the default workload, from the
    FETCH_SIZE / WRITE_SIZE counter runs summarised in the profile above (each
    counter scaled by what it reads for a device copy of known size); None for
    any other workload."""
    if args.method != "zb05" or args.particles != 1_000_000 or args.grid != 256:
        return None
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch = "):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


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
    ap.add_argument("--no-batch", action="store_true", help="skip the batched comparison")
    ap.add_argument("--batch-frames", type=int, default=64)
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (scale factor of the HBM "
                         "byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus %d is not supported "
                                   "(frames shard across GPUs without communication, see "
                                   "DESIGN.md)" % args.gpus}))
        sys.exit(2)

    import torch

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                          "copies": 5}))
        return

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
        roofline = {"bound": "mfma", "kernel": "k_field_zb05", "achieved": achieved_tflops,
                    "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                    "frac": achieved_tflops / (FP32_PEAK / 1e12),
                    "traffic": measured_traffic(args),
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
    del pos_k, dens_k, rec
    torch.cuda.empty_cache()
    if not args.no_batch:
        out["batched"] = batched_vs_per_frame(args.method, args.batch_frames, args.lanes,
                                              args.batch_reps, args.autotune)
        print(f"[bench] batched: {out['batched']}", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_cost()
        print(f"[bench] adjoint: {out['adjoint']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline and not yu13:
        out["cpu_baseline"] = cpu_baseline(make_frame(), args.method)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
