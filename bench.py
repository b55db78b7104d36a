"""Benchmark: full reconstruction frames/s on one MI355X.

Default workload (the headline `value`): a dam-break frame of 1,000,000
particles (a 160x80x80 lattice block at spacing 0.005 plus a ball of radius 0.1
at the same density, 10% of the particles with Gaussian noise, seeded subset),
cubic 256^3 scalar grid over the padded AABB, ZB05 field on the occupancy band
(the reference's ReconstructorGPUOursZB05 stage order).  A step = one whole
frame: spatial hashing -> surface particles -> band -> field -> marching cubes,
inputs resident in HBM (each step reads its own unsorted copy, since a frame
sorts its particles in place).

Beside the headline the same JSON line carries
  akinci12   the same frame with the Akinci12 colour-field band;
  yu13       the anisotropic (TC01) frame: same input, 384^3 grid;
  batched    64 random-shape frames, ZB05 and Yu13, batched vs per-frame;
  frame_sequence  the frame-sequence driver on .xyz copies of the headline
             frame, pipelined against strictly serial (read, reconstruct, write);
  adjoint    the 256,000-particle ball, backward vs forward time;
  each with its kernels' roofline figures, and the CPU oracle beside them.

The `roofline` object is the ZB05 field kernel's.  Its HBM traffic and L2 hit
rate come from rocprofv3 counter runs of this script in probe mode (child
processes), each HBM counter scaled by what it reads for a device copy of known
size in the same run.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("frames/s at 1M particles, 256^3 grid; field-stage HBM GB/s; batched vs per-frame")
HBM_PEAK = 8.0e12
FP32_PEAK = 157.3e12
COPY_MIB = 256


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ----------------------------------------------------------- algorithmic counts

def cell_counts(fr, ranges: np.ndarray) -> np.ndarray:
    """Particles per hash cell, shaped (rz, ry, rx)."""
    rx, ry, rz = fr.sp.res
    cnt = np.where(ranges[:, 0] == 0xFFFFFFFF, 0, ranges[:, 1].astype(np.int64) - ranges[:, 0])
    return cnt.reshape(rz, ry, rx)


class BlockSums:
    """Sums of a per-cell grid over clamped (2r+1)^3 cell blocks (the
    reference's candidate blocks), through a summed-area table."""

    def __init__(self, grid3: np.ndarray):
        rz, ry, rx = grid3.shape
        self.res = (rx, ry, rz)
        self.sat = np.zeros((rz + 1, ry + 1, rx + 1), np.int64)
        self.sat[1:, 1:, 1:] = grid3.astype(np.int64).cumsum(0).cumsum(1).cumsum(2)

    def __call__(self, cx, cy, cz, r: int) -> np.ndarray:
        (rx, ry, rz), s = self.res, self.sat
        x0, x1 = np.clip(cx - r, 0, rx - 1), np.clip(cx + r, 0, rx - 1) + 1
        y0, y1 = np.clip(cy - r, 0, ry - 1), np.clip(cy + r, 0, ry - 1) + 1
        z0, z1 = np.clip(cz - r, 0, rz - 1), np.clip(cz + r, 0, rz - 1) + 1
        return (s[z1, y1, x1] - s[z0, y1, x1] - s[z1, y0, x1] - s[z1, y1, x0]
                + s[z0, y0, x1] + s[z0, y1, x0] + s[z1, y0, x0] - s[z0, y0, x0])


def vertex_cells(fr, band_index: np.ndarray):
    """Hash cell (unclamped) of every band vertex, in the kernels' FP32
    arithmetic (vertex position by one rounding, then (p - min) / cell)."""
    sx, sy, _ = fr.sc.res
    v = band_index.astype(np.int64)
    idx = (v % sx, (v // sx) % sy, v // (sx * sy))
    out = []
    for a in range(3):
        vf = (idx[a].astype(np.float64) * float(fr.sc.cell) + float(fr.sc.min[a])).astype(np.float32)
        q = (vf - np.float32(fr.sp.min[a])) / np.float32(fr.sp.cell)
        out.append(np.trunc(np.clip(q, -4, fr.sp.res[a] + 4)).astype(np.int64))
    return out


def all_cells(fr):
    rx, ry, rz = fr.sp.res
    zz, yy, xx = np.meshgrid(np.arange(rz), np.arange(ry), np.arange(rx), indexing="ij")
    return xx.ravel(), yy.ravel(), zz.ravel()


def zb05_field_counts(fr, band_index, ranges):
    """P27 = sum over band vertices of the particles in the clamped 3^3 block of
    the vertex's hash cell; N_touched = particles in the union of those blocks."""
    c3 = cell_counts(fr, ranges)
    cx, cy, cz = vertex_cells(fr, band_index)
    p27 = int(BlockSums(c3)(cx, cy, cz, 1).sum())
    rx, ry, rz = fr.sp.res
    touched = np.zeros((rz, ry, rx), bool)
    uc = np.unique(np.stack([np.clip(cx, 0, rx - 1), np.clip(cy, 0, ry - 1),
                             np.clip(cz, 0, rz - 1)], 1), axis=0)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                touched[np.clip(uc[:, 2] + dz, 0, rz - 1), np.clip(uc[:, 1] + dy, 0, ry - 1),
                        np.clip(uc[:, 0] + dx, 0, rx - 1)] = True
    return p27, int(c3[touched].sum())


def colorfield_p27(fr, ranges):
    """Candidate pairs of the Akinci12 colour field: sum over particles of the
    particles in the clamped 3^3 block of their cell (self included)."""
    c3 = cell_counts(fr, ranges)
    cx, cy, cz = all_cells(fr)
    return int((c3.ravel() * BlockSums(c3)(cx, cy, cz, 1)).sum())


def yu13_counts(fr, band_index, ranges, involve):
    """Candidate pairs of the anisotropic path over clamped 5^3 blocks:
    smoothing P125(N), covariance P125(I), field P125 split into involved and
    other candidates; I = involved particles."""
    c3 = cell_counts(fr, ranges)
    inv3 = (involve.reshape(c3.shape) > 0)
    c_inv = np.where(inv3, c3, 0)
    c_non = c3 - c_inv
    bs_all, bs_inv, bs_non = BlockSums(c3), BlockSums(c_inv), BlockSums(c_non)
    cx, cy, cz = all_cells(fr)
    blk = bs_all(cx, cy, cz, 2)
    p_smooth = int((c3.ravel() * blk).sum())
    p_cov = int((c_inv.ravel() * blk).sum())
    vx, vy, vz = vertex_cells(fr, band_index)
    p_inv = int(bs_inv(vx, vy, vz, 2).sum())
    p_non = int(bs_non(vx, vy, vz, 2).sum())
    return dict(p125_smooth=p_smooth, p125_cov=p_cov, p125_field_inv=p_inv,
                p125_field_non=p_non, involved=int(c_inv.sum()))


# ------------------------------------------------------------------ frame runs

def make_frame(method: str, particles: int, grid: int):
    from fluidsurfacereconstructiongpu_amd import datasets as ds

    if method == "yu13":
        return ds.dam_break(n_target=particles, grid_res=grid, method="yu13", scale=1.3, iso=0.0)
    return ds.dam_break(n_target=particles, grid_res=grid, method=method)


def time_frames(method, fr, K, W, aniso=None):
    """K timed frames of one method after W warm-up frames, inputs resident;
    returns (reconstructor, total ms, per-frame stage ms (K, 6), stats)."""
    import torch

    from fluidsurfacereconstructiongpu_amd.reconstructor import Reconstructor, stats_dict

    dev = torch.device("cuda", 0)
    rec = Reconstructor(method, aniso=aniso)
    src_pos = torch.from_numpy(fr.pos).to(dev)
    src_dens = torch.from_numpy(fr.dens).to(dev)
    pos_k = [src_pos.clone() for _ in range(K + W)]
    dens_k = [src_dens.clone() for _ in range(K + W)]
    first = rec.run(fr, pos=src_pos.clone(), dens=src_dens.clone())
    cap = max(first.stats["num_mesh_vertices"] + 1024, 3 * 1024)
    sim = rec.params(fr)
    stream = torch.cuda.current_stream(dev)
    for i in range(W):
        rec.launch(fr, pos_k[i], dens_k[i], cap, sim=sim, stream=stream)
    torch.cuda.synchronize(dev)
    # the timed frames run without stage events (they cost ~3% of a frame);
    # the stage split comes from K more frames with them
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for i in range(W, W + K):
        rec.launch(fr, pos_k[i], dens_k[i], cap, sim=sim, stream=stream)
    ev1.record(stream)
    torch.cuda.synchronize(dev)
    total_ms = ev0.elapsed_time(ev1)
    for i in range(W, W + K):
        pos_k[i].copy_(src_pos)
        dens_k[i].copy_(src_dens)
    torch.cuda.synchronize(dev)
    rec.ctx.set_profiling(K)
    for i in range(W, W + K):
        rec.launch(fr, pos_k[i], dens_k[i], cap, sim=sim, stream=stream)
    torch.cuda.synchronize(dev)
    stages = rec.ctx.stage_times(K)
    rec.ctx.set_profiling(0)
    st = stats_dict(rec._stats)
    del pos_k, dens_k
    return rec, total_ms, stages, st


def stage_summary(stages):
    from fluidsurfacereconstructiongpu_amd import _lib

    frame_ms = stages.sum(axis=1)
    return {"frame_ms_p50": float(np.percentile(frame_ms, 50)),
            "frame_ms_p99": float(np.percentile(frame_ms, 99)),
            "stage_ms": {k: float(stages[:, i].mean()) for i, k in enumerate(_lib.STAGES)}}


def work_sums(rec, name):
    """Sums of a kernel's (staged, kept, third) work records: per wave
    (colour field) or per chunk up to the end marker (ZB05 field)."""
    _, cap = rec.ctx.buffer_ptr(name)
    w = rec.buffer(name, np.uint32, cap)
    end = np.nonzero(w == 0xFFFFFFFF)[0]
    if end.shape[0]:
        w = w[: end[0]]
    w = w[: w.shape[0] // 3 * 3].reshape(-1, 3).astype(np.int64)
    return [int(x) for x in w.sum(axis=0)]


def flop_roofline(kernel, flops, ms, **extra):
    achieved = flops / (ms * 1e-3) / 1e12
    # the pair tiles of these kernels run on the matrix cores (with VALU work
    # between the products); the colour field is VALU code
    mfma = kernel in ("k_field_zb05", "k_field_aniso", "k_smooth + k_covariance")
    out = {"bound": "mfma" if mfma else "valu", "kernel": kernel,
           "achieved": achieved, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
           "frac": achieved / (FP32_PEAK / 1e12), "flops_per_launch": float(flops),
           "kernel_ms": ms}
    out.update(extra)
    return out


def zb05_roofline(rec, fr, st, stage):
    B = st["num_band_vertices"]
    band_index = rec.buffer("band_index", np.uint32, B)
    ranges = rec.buffer("cell_ranges", np.uint32, 2 * fr.sp.count).reshape(-1, 2)
    p27, n_touched = zb05_field_counts(fr, band_index, ranges)
    staged, kept, steps = work_sums(rec, "field_work")
    ms = float(stage["stage_ms"]["field"])
    pairs_exec = 256 * steps
    alg_bytes = 8.0 * B + 8.0 * fr.sp.count + 12.0 * n_touched
    ref_rate = 20.0 * p27 / (ms * 1e-3) / 1e12
    return flop_roofline(
        "k_field_zb05", 20.0 * pairs_exec, ms,
        executed_pairs=pairs_exec, candidates_staged=staged, candidates_kept=kept,
        pairs_p27=p27, reference_pair_rate=ref_rate,
        reference_pair_frac=ref_rate / (FP32_PEAK / 1e12),
        note=("achieved = 20 FLOP per pair the kernel evaluates (256 per 16x16 "
              "vertex/candidate tile left after its box cull, counted on the device) / "
              "kernel time; reference_pair_rate counts 20 FLOP per candidate pair of the "
              "reference (P27: the clamped 3^3 cell block of each band vertex), most of "
              "which the cull never evaluates"),
        algorithmic_bytes=alg_bytes, algorithmic_gbps=alg_bytes / (ms * 1e-3) / 1e9,
        traffic=None)


def akinci_line(fr, K, W):
    rec, total_ms, stages, st = time_frames("akinci12", fr, K, W)
    stage = stage_summary(stages)
    ranges = rec.buffer("cell_ranges", np.uint32, 2 * fr.sp.count).reshape(-1, 2)
    p27 = colorfield_p27(fr, ranges)
    staged, kept, inh = work_sums(rec, "color_work")
    ms = float(stage["stage_ms"]["surface"])
    # 8 FLOP per candidate (difference, |r|^2, test), 12 more within h (square
    # root, kernel, gradient accumulation)
    colour = flop_roofline("k_colorfield", 8.0 * p27 + 12.0 * inh, ms, pairs_p27=p27,
                           pairs_within_h=inh, candidates_kept=kept,
                           note="time = the surface stage (k_colorfield and its chunk list)")
    field = zb05_roofline(rec, fr, st, stage)
    out = {"value": K / (total_ms * 1e-3), "unit": "frames/s",
           "workload": "dam-break frame, %d particles, 256^3 grid, Akinci12 colour-field band, "
                       "ZB05 field, full marching cubes" % fr.n,
           **stage, "counts": counts_of(st), "roofline_colorfield": colour,
           "roofline_field": field}
    return out, rec, st


def aniso_work(rec):
    """(staged, kept) candidate totals of the last smoothing, covariance and
    field launches, from their per-wave records."""
    _, cap = rec.ctx.buffer_ptr("aniso_work")
    w = rec.buffer("aniso_work", np.uint32, cap).astype(np.int64)
    per = cap // 3
    out = []
    for k in range(3):
        sec = w[k * per:(k + 1) * per]
        waves = int(sec[0])
        rec2 = sec[1:1 + 2 * waves].reshape(-1, 2)
        out.append((int(rec2[:, 0].sum()), int(rec2[:, 1].sum())))
    return out


# FP32 operations per evaluated pair (fused multiply-add = 2), from the pair
# loops: smoothing (distance by expansion, square root, weight, 4 sums),
# covariance (difference, distance, count, weight, 7 sums), field (9 FMA
# quadratic form, clamp, cube, weighted sum).
ANISO_PAIR_FLOP = {"smooth": 19.0, "cov": 33.0, "field": 24.0}


def yu13_line(fr, aniso, K, W):
    rec, total_ms, stages, st = time_frames("yu13", fr, K, W, aniso=aniso)
    stage = stage_summary(stages)
    B = st["num_band_vertices"]
    band_index = rec.buffer("band_index", np.uint32, B)
    ranges = rec.buffer("cell_ranges", np.uint32, 2 * fr.sp.count).reshape(-1, 2)
    involve = rec.buffer("involve_counts", np.uint32, fr.sp.count)
    c = yu13_counts(fr, band_index, ranges, involve)
    (sm_st, sm_kept), (cv_st, cv_kept), (fd_st, fd_kept) = aniso_work(rec)
    ms_field = float(stage["stage_ms"]["field"])
    ms_prep = float(stage["stage_ms"]["prep"])
    f_field = ANISO_PAIR_FLOP["field"] * 64.0 * fd_kept
    f_prep = 64.0 * (ANISO_PAIR_FLOP["smooth"] * sm_kept + ANISO_PAIR_FLOP["cov"] * cv_kept)
    ref_field = 30.0 * c["p125_field_inv"] + 15.0 * c["p125_field_non"]
    ref_prep = 20.0 * c["p125_smooth"] + 32.0 * c["p125_cov"] + 600.0 * c["involved"]
    field = flop_roofline(
        "k_field_aniso", f_field, ms_field, candidates_staged=fd_st, candidates_kept=fd_kept,
        executed_pairs=64 * fd_kept, pairs_involved=c["p125_field_inv"],
        pairs_other=c["p125_field_non"],
        reference_pair_rate=ref_field / (ms_field * 1e-3) / 1e12,
        note=("achieved = 24 FLOP per pair evaluated (64 lanes x each candidate kept by the "
              "chunk-box cull, counted on the device) / field stage time; particles of "
              "non-involved cells enter as 35 per-block moments (FP64 polynomial per vertex), "
              "not as pairs; reference_pair_rate counts 30 P125(involved) + 15 P125(other) "
              "FLOP of the reference's 5^3-block loop"))
    prep = flop_roofline(
        "k_smooth + k_covariance", f_prep, ms_prep,
        smooth_staged=sm_st, smooth_kept=sm_kept, cov_staged=cv_st, cov_kept=cv_kept,
        pairs_smooth=c["p125_smooth"], pairs_cov=c["p125_cov"], involved=c["involved"],
        reference_pair_rate=ref_prep / (ms_prep * 1e-3) / 1e12,
        note=("achieved = 64 lanes x (19 FLOP per kept smoothing candidate + 33 per kept "
              "covariance candidate) / prep stage time (the two kernels, their chunk lists "
              "and the involved list); reference_pair_rate counts 20 P125(N) + 32 P125(I) + "
              "600 I"))
    out = {"value": K / (total_ms * 1e-3), "unit": "frames/s",
           "workload": "dam-break frame, %d particles, 384^3 grid, Yu13 anisotropic (TC01) "
                       "field, lambda %.2f kn %.2f kr %.1f Ne %d, full marching cubes"
                       % (fr.n, aniso["lam"], aniso["kn"], aniso["kr"], aniso["ne"]),
           **stage, "counts": counts_of(st, yu13=True), "roofline_field": field,
           "roofline_prep": prep}
    return out, rec, st


def counts_of(st, yu13=False):
    keys = ["num_surface_particles", "num_band_vertices", "num_valid_cubes", "num_mesh_vertices"]
    if yu13:
        keys.insert(1, "num_involved_particles")
    return {k: st[k] for k in keys}


# ------------------------------------------------------------- batched / adjoint

def batched_vs_per_frame(method: str, frames: int, lanes: int, reps: int, autotune: bool):
    """The batched workload: `frames` random-shape frames (particle counts
    log-uniform in [2e4, 2e6], 96^3-320^3 grids), each starting from its own
    resident unsorted copy.  Per-frame path: the frames one after the other on
    one stream (copy + frame each).  Batched path: the same work through
    fsr_batch_run (one graph launch per pass)."""
    import ctypes

    import torch

    from fluidsurfacereconstructiongpu_amd import _lib
    from fluidsurfacereconstructiongpu_amd import datasets as ds
    from fluidsurfacereconstructiongpu_amd.reconstructor import METHODS, Batch

    kw = dict(method="yu13", iso=0.0, scale=1.3) if method == "yu13" else {}
    t0 = time.perf_counter()
    frs = [ds.random_shapes_frame(i, **kw) for i in range(frames)]
    log(f"batch workload ({method}): {frames} frames, {sum(f.n for f in frs)} particles, "
        f"generated in {time.perf_counter() - t0:.1f} s")
    aniso = ds.yu13_params(frs[0], True) if method == "yu13" else None
    batch = Batch(method, frs, lanes=lanes, aniso=aniso)
    dev = batch.dev
    stream = torch.cuda.current_stream(dev)
    tuned = None
    if autotune:
        # the lanes' field-kernel workgroups per CU, timed on the whole batch
        choice = ctypes.c_uint32()
        ap = batch.rec.aniso_param()
        _lib.check(batch.ctx.lib.fsr_batch_autotune(
            batch.handle, METHODS[method], ctypes.byref(ap), batch.descs, len(frs), 3,
            ctypes.byref(choice), ctypes.c_void_p(stream.cuda_stream)))
        torch.cuda.synchronize(dev)
        tuned = int(choice.value)
    lib, ctx = batch.ctx.lib, batch.ctx.handle
    ap = batch.rec.aniso_param()

    def per_frame_pass(events=None, sync=False):
        for i, d in enumerate(batch.descs):
            b = batch.bufs[i]
            if events is not None:
                events[i][0].record(stream)
            b["pos"].copy_(b["src_pos"])
            b["dens"].copy_(b["src_dens"])
            _lib.check(lib.fsr_reconstruct_frame(
                ctx, METHODS[method], ctypes.byref(d.sim), ctypes.byref(ap),
                ctypes.byref(d.sp_grid), ctypes.byref(d.sc_grid), d.pos, d.dens, d.num_particles,
                d.phi, d.tri_pos, d.tri_nrm, d.tri_capacity, d.stats,
                ctypes.c_void_p(stream.cuda_stream)))
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
    t_sync = time.perf_counter()
    for r in range(reps):
        per_frame_pass(sync=True)
    torch.cuda.synchronize(dev)
    sync_ms = (time.perf_counter() - t_sync) * 1e3 / reps
    # the frames' neighbour-kernel time (field stage, and for Yu13 the prep
    # stage: smoothing and covariance), from one more per-frame pass with stage
    # events: what a batch has left to overlap is the rest of the per-frame time
    # (and the idle tails of those persistent kernels on small frames)
    n_frames = len(frs)
    batch.ctx.set_profiling(n_frames)
    per_frame_pass()
    torch.cuda.synchronize(dev)
    stages = batch.ctx.stage_times(n_frames)
    batch.ctx.set_profiling(0)
    heavy_ms = float(stages[:, 4].sum() + (stages[:, 3].sum() if method == "yu13" else 0.0))
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
        "lanes": lanes, "passes": reps,
        "autotuned_lane_field_workgroups_per_cu": tuned,
        "batched_frames_per_s": n / (bat_ms * 1e-3),
        "per_frame_sync_frames_per_s": n / (sync_ms * 1e-3),
        "per_frame_async_frames_per_s": n / (per_ms * 1e-3),
        "speedup_vs_per_frame_sync": sync_ms / bat_ms,
        "speedup_vs_per_frame_async": per_ms / bat_ms,
        "neighbour_kernel_ms_sum": heavy_ms,
        "neighbour_kernel_share_of_per_frame_sync": heavy_ms / sync_ms,
        "batched_frame_ms_p50": float(np.percentile(bat_frame_ms, 50)),
        "batched_frame_ms_p99": float(np.percentile(bat_frame_ms, 99)),
        "per_frame_ms_p50": float(np.percentile(per_frame_ms, 50)),
        "per_frame_ms_p99": float(np.percentile(per_frame_ms, 99)),
    }


def frame_sequence(fr, frames: int = 6):
    """Frame-sequence driver (the reference's ReconstructorGPU::reconstruct flow:
    read .xyz, reconstruct, write .obj, frame after frame) on `frames` copies of
    the headline frame written as .xyz files: end-to-end frames/s of the
    pipelined order (frame k+1 parsed and uploaded while frame k runs and its
    mesh is written) against --serial (strictly one after another)."""
    import re

    from fluidsurfacereconstructiongpu_amd import frame_io
    res = {"frames": frames}
    with tempfile.TemporaryDirectory(prefix="fsr_seq_") as d:
        lo, hi = fr.pos.min(axis=0), fr.pos.max(axis=0)
        first = os.path.join(d, "frame_000000.xyz")
        frame_io.write_xyz(first, lo, hi, fr.h, fr.particle_radius, fr.mass, fr.pos, fr.dens)
        for k in range(1, frames):
            os.link(first, os.path.join(d, "frame_%06d.xyz" % k))
        for mode in ("serial", "pipelined"):
            cmd = [frame_io.DRIVER, d + "/", "frame_%06d", "0", str(frames)]
            if mode == "serial":
                cmd.append("--serial")
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
            if r.returncode != 0:
                raise RuntimeError("fsr_reconstruct failed: %s" % r.stderr[-500:])
            m = re.search(r"(\d+) frame\(s\) in ([0-9.eE+-]+) ms end to end", r.stdout)
            rec = [float(x) for x in re.findall(r"took ([0-9.eE+-]+) milliseconds", r.stdout)]
            res[mode + "_frames_per_s"] = int(m.group(1)) * 1e3 / float(m.group(2))
            res[mode + "_reconstruct_ms_median"] = float(np.median(rec))
            g = re.search(r"Scalar Field Resolution: (\d+) \* (\d+) \* (\d+)", r.stdout)
            res["grid"] = [int(g.group(i)) for i in (1, 2, 3)]
            for k in range(frames):
                obj = os.path.join(d, "frame_%06d.obj" % k)
                if os.path.exists(obj):
                    res.setdefault("obj_bytes", os.path.getsize(obj))
                    os.remove(obj)
    res["speedup"] = res["pipelined_frames_per_s"] / res["serial_frames_per_s"]
    res["workload"] = ("%d .xyz frames of %d particles (grid from the reference's rule), ZB05, "
                       "soup .obj written per frame" % (frames, fr.n))
    return res


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
    from fluidsurfacereconstructiongpu_amd.reconstructor import stats_dict

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
    st = stats_dict(sf.rec._stats)
    res = {"workload": "jittered lattice ball, %d particles, 192^3 grid, ZB05, sphere loss "
                       "over grid and soup" % fr.n,
           "t_fwd_ms": t_fwd, "t_bwd_ms": t_bwd, "t_bwd_field_ms": t_field,
           "t_bwd_vertex_term_ms": t_bwd - t_field, "ratio_bwd_fwd": t_bwd / t_fwd,
           "band_vertices": st["num_band_vertices"], "soup_vertices": st["num_mesh_vertices"],
           "reps": reps}
    sf.close()
    return res


# ------------------------------------------------------------ counter passes

def counter_probe():
    """Probe mode (run under rocprofv3 by counter_pass): one device copy of
    COPY_MIB MiB (the counters' calibration), then the default ZB05 frame
    (1 sizing run, 1 warm-up, 3 frames)."""
    import torch

    nbytes = COPY_MIB << 20
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    fr = make_frame("zb05", 1_000_000, 256)
    time_frames("zb05", fr, 3, 1)


def _norm(k: str) -> str:
    return "".join(ch for ch in k.lower() if ch.isalnum())


def counter_pass(counters, timeout_s=240):
    """Runs this script in probe mode under `rocprofv3 --pmc counters` and
    returns {kernel name: {counter: [per-dispatch values]}}."""
    exe = shutil.which("rocprofv3")
    if not exe:
        raise RuntimeError("rocprofv3 not found")
    out = tempfile.mkdtemp(prefix="fsr_pmc_", dir="/tmp")
    cmd = [exe, "--pmc", *counters, "-d", out, "-o", "run", "--output-format", "csv", "--",
           sys.executable, os.path.abspath(__file__), "--counter-probe"]
    env = dict(os.environ, TMPDIR="/tmp")
    p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                         stderr=subprocess.PIPE, start_new_session=True)
    try:
        _, err = p.communicate(timeout=timeout_s)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.communicate()
        raise RuntimeError("counter pass %s timed out" % counters)
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError("counter pass %s (rc %d) wrote no counter file: %s"
                           % (counters, p.returncode, err.decode(errors="replace")[-400:]))
    per = {}
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                r = {_norm(k): v for k, v in row.items() if k}
                name = r.get("kernelname", "")
                disp = r.get("dispatchid") or r.get("correlationid")
                cname, val = r.get("countername"), r.get("countervalue")
                if cname is None or val is None:
                    continue
                d = per.setdefault(name, {}).setdefault(cname, {})
                d[disp] = d.get(disp, 0.0) + float(val)
    shutil.rmtree(out, ignore_errors=True)
    return {k: {c: list(v.values()) for c, v in cs.items()} for k, cs in per.items()}


def _kernel_values(per, kernel, counter):
    import re

    pat = re.compile(r"\b" + kernel + r"\b(?!_)")
    vals = []
    for name, cs in per.items():
        if pat.search(name) and counter in cs:
            vals += cs[counter]
    return vals


def field_traffic():
    """HBM bytes per k_field_zb05 launch (FETCH_SIZE and WRITE_SIZE runs, each
    scaled by its reading for the calibration copy) and the kernel's L2 hit
    rate (TCC_HIT / (TCC_HIT + TCC_MISS))."""
    copy_bytes = float(COPY_MIB << 20)
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        per = counter_pass([counter])
        field = _kernel_values(per, "k_field_zb05", counter)
        copies = []
        for name, cs in per.items():
            if "copyBuffer" in name and counter in cs:
                copies += cs[counter]
        if not field or not copies:
            raise RuntimeError("%s: kernel or calibration copy missing from the counters" % counter)
        cal = max(copies)          # the COPY_MIB copy is the largest copy of the run
        scale = copy_bytes / (cal * 1024.0)
        res[counter] = dict(kib=float(np.mean(field)), cal_kib=cal, scale=scale,
                            bytes=float(np.mean(field)) * 1024.0 * scale)
    per = counter_pass(["TCC_HIT_sum", "TCC_MISS_sum"])
    hit = _kernel_values(per, "k_field_zb05", "TCC_HIT_sum")
    miss = _kernel_values(per, "k_field_zb05", "TCC_MISS_sum")
    l2 = float(np.sum(hit) / max(1.0, np.sum(hit) + np.sum(miss))) if hit else None
    return dict(read=res["FETCH_SIZE"]["bytes"], written=res["WRITE_SIZE"]["bytes"],
                fetch_scale=res["FETCH_SIZE"]["scale"], write_scale=res["WRITE_SIZE"]["scale"],
                l2_hit_rate=l2)


# -------------------------------------------------------------- CPU baselines

def cpu_baseline(fr, method, gpu_stats, aniso=None, budget_s=25.0):
    """The CPU oracle (FP64, OpenMP on every host core) on the same frame, timed
    over whole frames (median of up to 3 within budget_s).  The frame's counts
    are checked against the GPU's: discrete stages (surface particles, involved
    particles, band vertices) exactly; valid cubes and soup vertices, which
    depend on field values within 1e-5 of the iso value, to 1e-4 relative."""
    import oracle

    threads = oracle.threads()
    times, counts = [], None
    t_total = 0.0
    while len(times) < 3 and t_total < budget_s:
        t0 = time.perf_counter()
        counts = oracle.frame_yu13(fr, aniso) if method == "yu13" else oracle.frame(method, fr)[0]
        dt = time.perf_counter() - t0
        times.append(dt)
        t_total += dt
        if dt > 10.0:
            break
    med = float(np.median(times))
    exact = [("num_surface_particles", "surface_particles"),
             ("num_band_vertices", "band_vertices")]
    if method == "yu13":
        exact.append(("num_involved_particles", "involved_particles"))
    for g, c in exact:
        if gpu_stats[g] != counts[c]:
            raise AssertionError(f"{method}: GPU {g} = {gpu_stats[g]} but the oracle counts "
                                 f"{counts[c]}")
    for g, c in (("num_valid_cubes", "valid_cubes"), ("num_mesh_vertices", "mesh_vertices")):
        if abs(gpu_stats[g] - counts[c]) > 1e-4 * max(1, counts[c]):
            raise AssertionError(f"{method}: GPU {g} = {gpu_stats[g]}, oracle {counts[c]}")
    return {"value": 1.0 / med, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} full frame(s) of the same workload through the FP64 "
                      f"oracle (OpenMP), median {med:.2f} s",
            "nproc": os.cpu_count(),
            "counts_checked": {c: [gpu_stats[g], counts[c]] for g, c in
                               exact + [("num_valid_cubes", "valid_cubes"),
                                        ("num_mesh_vertices", "mesh_vertices")]}}


# ----------------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-akinci", action="store_true", help="skip the Akinci12 line")
    ap.add_argument("--no-yu13", action="store_true", help="skip the Yu13 line")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched comparison")
    ap.add_argument("--batch-frames", type=int, default=64)
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default lane launch shape (no fsr_batch_autotune)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost")
    ap.add_argument("--no-sequence", action="store_true",
                    help="skip the frame-sequence driver (pipelined vs serial)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter runs (traffic, L2 hit rate)")
    ap.add_argument("--counter-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--only", choices=["zb05", "akinci12", "yu13", "adjoint", "batch",
                                       "sequence"],
                    help="run one workload alone (for kernel traces) and print its result "
                         "instead of the benchmark line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus %d is not supported "
                                   "(frames shard across GPUs without communication, see "
                                   "DESIGN.md)" % args.gpus}))
        sys.exit(2)
    if args.counter_probe:
        counter_probe()
        return

    import torch

    from fluidsurfacereconstructiongpu_amd import datasets as ds

    K, W = args.steps, args.warmup
    if args.only:
        if args.only in ("zb05", "akinci12"):
            fr = make_frame(args.only, args.particles, args.grid)
            if args.only == "zb05":
                rec, total_ms, stages, st = time_frames("zb05", fr, K, W)
                res = {"value": K / (total_ms * 1e-3), "unit": "frames/s",
                       **stage_summary(stages), "counts": counts_of(st)}
            else:
                res = akinci_line(fr, K, W)[0]
        elif args.only == "yu13":
            fr = make_frame("yu13", args.particles, 384)
            res = yu13_line(fr, ds.yu13_params(fr, True), K, W)[0]
        elif args.only == "adjoint":
            res = adjoint_cost(reps=K)
        elif args.only == "sequence":
            res = frame_sequence(make_frame("zb05", args.particles, args.grid))
        else:
            res = {m: batched_vs_per_frame(m, args.batch_frames, args.lanes, args.batch_reps,
                                           not args.no_autotune) for m in ("zb05", "yu13")}
        print(json.dumps({"only": args.only, "steps": K, "warmup": W, "result": res}),
              flush=True)
        return
    t0 = time.perf_counter()
    fr = make_frame("zb05", args.particles, args.grid)
    log(f"workload generated in {time.perf_counter() - t0:.1f} s: N={fr.n} sp={fr.sp.res} "
        f"sc={fr.sc.res}")
    rec, total_ms, stages, st = time_frames("zb05", fr, K, W)
    log(f"{K} frames in {total_ms:.2f} ms; stats {st}")
    stage = stage_summary(stages)
    roofline = zb05_roofline(rec, fr, st, stage)
    del rec
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
        "config": {"workload": "dam-break frame, %d particles, %d^3 scalar grid, ZB05 field on "
                               "the occupancy band, full marching cubes" % (fr.n, args.grid),
                   "particles": fr.n, "grid": list(fr.sc.res), "hash_grid": list(fr.sp.res),
                   "method": "zb05", "parallelism": "single GPU"},
        "roofline": roofline,
        **stage,
        "counts": counts_of(st),
    }
    torch.cuda.empty_cache()
    zb05_stats = st
    akinci_stats = yu13_stats = None
    if not args.no_akinci:
        out["akinci12"], rec, akinci_stats = akinci_line(fr, K, W)
        del rec
        log(f"akinci12: {out['akinci12']['value']:.0f} frames/s")
        torch.cuda.empty_cache()
    fr3 = aniso = None
    if not args.no_yu13:
        fr3 = make_frame("yu13", args.particles, 384)
        aniso = ds.yu13_params(fr3, True)
        out["yu13"], rec, yu13_stats = yu13_line(fr3, aniso, max(5, K // 3), 2)
        del rec
        log(f"yu13: {out['yu13']['value']:.0f} frames/s")
        torch.cuda.empty_cache()
    if not args.no_batch:
        out["batched"] = batched_vs_per_frame("zb05", args.batch_frames, args.lanes,
                                              args.batch_reps, not args.no_autotune)
        log(f"batched zb05: {out['batched']}")
        torch.cuda.empty_cache()
        out["batched_yu13"] = batched_vs_per_frame("yu13", args.batch_frames, args.lanes,
                                                   args.batch_reps, not args.no_autotune)
        log(f"batched yu13: {out['batched_yu13']}")
        torch.cuda.empty_cache()
    if not args.no_sequence:
        try:
            out["frame_sequence"] = frame_sequence(fr)
            log(f"frame sequence: {out['frame_sequence']}")
        except Exception as e:  # the headline line stands without it
            out["frame_sequence"] = {"error": str(e)}
            log(f"frame sequence failed: {e}")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_cost()
        log(f"adjoint: {out['adjoint']}")
    if not args.no_counters:
        try:
            tr = field_traffic()
            roofline["traffic"] = tr["read"] + tr["written"]
            roofline["traffic_read"] = tr["read"]
            roofline["traffic_written"] = tr["written"]
            roofline["traffic_scale"] = [tr["fetch_scale"], tr["write_scale"]]
            roofline["l2_hit_rate"] = tr["l2_hit_rate"]
            roofline["traffic_gbps"] = roofline["traffic"] / (roofline["kernel_ms"] * 1e-3) / 1e9
            log(f"traffic: {tr}")
        except Exception as e:  # the line is still valid without the counter figures
            roofline["traffic_note"] = "counter runs failed: %s" % e
            log(roofline["traffic_note"])
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(fr, "zb05", zb05_stats)
        log(f"cpu baseline zb05: {out['cpu_baseline']}")
        if akinci_stats is not None:
            out["akinci12"]["cpu_baseline"] = cpu_baseline(fr, "akinci12", akinci_stats)
        if yu13_stats is not None:
            out["yu13"]["cpu_baseline"] = cpu_baseline(fr3, "yu13", yu13_stats, aniso=aniso,
                                                       budget_s=1.0)
            log(f"cpu baseline yu13: {out['yu13']['cpu_baseline']}")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
