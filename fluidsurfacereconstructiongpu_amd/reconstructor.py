"""Host-side wrapper over the C-ABI for Python callers (tests, bench, autograd).

Mirrors the reference's per-frame flow (GPU/ReconstructorGPU.cpp:25-79): a frame's
particles go in, the scalar grid and the triangle soup come out.  Device memory
comes from torch tensors; every computation runs in libfsr.so.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _lib
from .datasets import Frame, Grid, sim_param_fields

METHODS = {"zb05": _lib.METHOD_ZB05, "akinci12": _lib.METHOD_AKINCI12, "yu13": _lib.METHOD_YU13}


def grid_info(g: Grid) -> _lib.GridInfo:
    gi = _lib.GridInfo()
    gi.min_pos = _lib.Float3(float(g.min[0]), float(g.min[1]), float(g.min[2]))
    gi.max_pos = _lib.Float3(float(g.max[0]), float(g.max[1]), float(g.max[2]))
    gi.cell_size = float(g.cell)
    gi.resolution = _lib.UInt3(int(g.res[0]), int(g.res[1]), int(g.res[2]))
    return gi


def sim_param(fields: dict) -> _lib.SimParam:
    sp = _lib.SimParam()
    for k, v in fields.items():
        setattr(sp, k, v)
    return sp


class Context:
    """Owns one fsr_context (scratch arena) on a HIP device."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.handle = ctypes.c_void_p()
        _lib.check(self.lib.fsr_create(ctypes.byref(self.handle), device))
        self.device = device

    def close(self):
        if self.handle:
            self.lib.fsr_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_surface_mode(self, radius_count: bool) -> None:
        """Surface particles of ZB05 / Akinci12 frames: the method's own test
        (False, the reference's) or fewer than 25 particles closer than the
        smoothing radius (True)."""
        _lib.check(self.lib.fsr_set_surface_mode(self.handle, 1 if radius_count else 0))

    def set_profiling(self, ring_frames: int) -> None:
        _lib.check(self.lib.fsr_set_profiling(self.handle, int(ring_frames)))

    def stage_times(self, max_frames: int) -> np.ndarray:
        """(frames, 6) stage durations in ms of the frames recorded since the
        last call (blocking)."""
        out = (ctypes.c_float * (max_frames * len(_lib.STAGES)))()
        n = ctypes.c_uint32()
        _lib.check(self.lib.fsr_stage_times(self.handle, out, max_frames, ctypes.byref(n)))
        a = np.frombuffer(out, dtype=np.float32).reshape(max_frames, len(_lib.STAGES))
        return a[: n.value].copy()

    def buffer_ptr(self, name: str):
        p = ctypes.c_void_p()
        n = ctypes.c_uint64()
        _lib.check(self.lib.fsr_frame_buffer(self.handle, _lib.BUF[name], ctypes.byref(p),
                                             ctypes.byref(n)))
        return p.value, n.value


@dataclasses.dataclass
class FrameResult:
    phi: object            # torch float32 (V,)
    tri_pos: object        # torch float32 (capacity, 3)
    tri_nrm: object
    stats: dict

    @property
    def num_mesh_vertices(self) -> int:
        return int(self.stats["num_mesh_vertices"])

    def soup(self):
        n = self.num_mesh_vertices
        return self.tri_pos[:n], self.tri_nrm[:n]


def stats_dict(t) -> dict:
    a = t.cpu().numpy().view(np.uint32)
    names = [f[0] for f in _lib.FrameStats._fields_ if f[0] != "reserved"]
    return {k: int(a[i]) for i, k in enumerate(names)}


class Reconstructor:
    """One method (zb05 / akinci12 / yu13) over frames that live on the GPU."""

    def __init__(self, method: str = "zb05", device: int = 0, ctx: Context | None = None,
                 aniso: dict | None = None):
        """aniso (Yu13 only): dict(lam, kn, kr, ks, ne); default the reference's
        constants (lambda 0.5, kn 1, kr 4, Ne 25)."""
        import torch

        self.torch = torch
        self.method = method
        self.ctx = ctx or Context(device)
        self.device = torch.device("cuda", device)
        self.aniso = dict(lam=0.5, kn=1.0, kr=4.0, ks=1.0, ne=25)
        if aniso:
            self.aniso.update(aniso)
        self._phi = None
        self._tri = None
        self._nrm = None
        self._stats = torch.zeros(_lib.STATS_WORDS, dtype=torch.int32, device=self.device)

    def params(self, fr: Frame, **kw) -> _lib.SimParam:
        if self.method == "yu13":
            kw.setdefault("ratio", 3)
            kw.setdefault("scale", 1.3)
            kw.setdefault("lam", self.aniso["lam"])
            kw.setdefault("min_neighbors", self.aniso["ne"])
        ratio = kw.pop("ratio", 2)
        return sim_param(sim_param_fields(fr, ratio=ratio, **kw))

    def aniso_param(self) -> _lib.AnisoParam:
        a = _lib.AnisoParam()
        a.kr, a.ks, a.kn = self.aniso["kr"], self.aniso["ks"], self.aniso["kn"]
        return a

    def _ensure(self, V: int, cap: int):
        torch = self.torch
        if self._phi is None or self._phi.numel() < V:
            self._phi = torch.empty(V, dtype=torch.float32, device=self.device)
        if self._tri is None or self._tri.shape[0] < cap:
            self._tri = torch.empty((cap, 3), dtype=torch.float32, device=self.device)
            self._nrm = torch.empty((cap, 3), dtype=torch.float32, device=self.device)

    def launch(self, fr: Frame, pos, dens, tri_capacity: int, sim=None, stream=None):
        """Enqueues one frame; pos (N, 3) and dens (N,) are device tensors that
        are sorted in place.  Returns without synchronising."""
        torch = self.torch
        V = fr.sc.count
        self._ensure(V, max(1, tri_capacity))
        sim = sim or self.params(fr)
        ap = self.aniso_param()
        spg, scg = grid_info(fr.sp), grid_info(fr.sc)
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        st = self.ctx.lib.fsr_reconstruct_frame(
            self.ctx.handle, METHODS[self.method], ctypes.byref(sim), ctypes.byref(ap),
            ctypes.byref(spg),
            ctypes.byref(scg), ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(dens.data_ptr()),
            pos.shape[0], ctypes.c_void_p(self._phi.data_ptr()),
            ctypes.c_void_p(self._tri.data_ptr()), ctypes.c_void_p(self._nrm.data_ptr()),
            tri_capacity, ctypes.c_void_p(self._stats.data_ptr()), ctypes.c_void_p(s.cuda_stream))
        _lib.check(st)

    def run(self, fr: Frame, tri_capacity: int | None = None, pos=None, dens=None) -> FrameResult:
        """Uploads (unless device tensors are given), runs one frame, synchronises,
        and grows the soup buffer and reruns once if it overflowed."""
        torch = self.torch
        if pos is None:
            pos = torch.from_numpy(fr.pos).to(self.device)
            dens = torch.from_numpy(fr.dens).to(self.device)
        cap = tri_capacity or max(3 * 1024, 6 * fr.n)
        p0, d0 = pos.clone(), dens.clone()
        self.launch(fr, pos, dens, cap)
        torch.cuda.synchronize(self.device)
        stats = stats_dict(self._stats)
        if stats["overflow"]:
            cap = stats["num_mesh_vertices"]
            pos.copy_(p0)
            dens.copy_(d0)
            self.launch(fr, pos, dens, cap)
            torch.cuda.synchronize(self.device)
            stats = stats_dict(self._stats)
        self.pos_sorted, self.dens_sorted = pos, dens
        V = fr.sc.count
        return FrameResult(self._phi[:V], self._tri, self._nrm, stats)

    def run_indexed(self, fr: Frame, vtx_capacity: int | None = None,
                    idx_capacity: int | None = None, pos=None, dens=None):
        """One frame with the deduplicated mesh (fsr_reconstruct_frame_indexed);
        synchronises, grows the mesh buffers and reruns once on overflow.
        Returns (phi, vertices (U, 3), normals (U, 3), indices (3T,), stats)."""
        torch = self.torch
        if pos is None:
            pos = torch.from_numpy(fr.pos).to(self.device)
            dens = torch.from_numpy(fr.dens).to(self.device)
        vcap = vtx_capacity or max(1024, 2 * fr.n)
        icap = idx_capacity or max(3 * 1024, 6 * fr.n)
        p0, d0 = pos.clone(), dens.clone()
        V = fr.sc.count
        self._ensure(V, 1)
        sim = self.params(fr)
        ap = self.aniso_param()
        spg, scg = grid_info(fr.sp), grid_info(fr.sc)
        for attempt in range(2):
            vtx = torch.empty((max(vcap, 1), 3), dtype=torch.float32, device=self.device)
            nrm = torch.empty((max(vcap, 1), 3), dtype=torch.float32, device=self.device)
            idx = torch.empty(max(icap, 1), dtype=torch.int32, device=self.device)
            s = torch.cuda.current_stream(self.device)
            _lib.check(self.ctx.lib.fsr_reconstruct_frame_indexed(
                self.ctx.handle, METHODS[self.method], ctypes.byref(sim), ctypes.byref(ap),
                ctypes.byref(spg), ctypes.byref(scg), ctypes.c_void_p(pos.data_ptr()),
                ctypes.c_void_p(dens.data_ptr()), pos.shape[0],
                ctypes.c_void_p(self._phi.data_ptr()), ctypes.c_void_p(vtx.data_ptr()),
                ctypes.c_void_p(nrm.data_ptr()), vcap, ctypes.c_void_p(idx.data_ptr()), icap,
                ctypes.c_void_p(self._stats.data_ptr()), ctypes.c_void_p(s.cuda_stream)))
            torch.cuda.synchronize(self.device)
            stats = stats_dict(self._stats)
            if not stats["overflow"] or attempt == 1:
                break
            vcap = stats["num_unique_vertices"]
            icap = stats["num_mesh_vertices"]
            pos.copy_(p0)
            dens.copy_(d0)
        self.pos_sorted, self.dens_sorted = pos, dens
        U, T3 = stats["num_unique_vertices"], stats["num_mesh_vertices"]
        return self._phi[:V], vtx[:U], nrm[:U], idx[:T3], stats

    def buffer(self, name: str, dtype, count: int):
        """Copy of an intermediate buffer of the last frame (numpy)."""
        torch = self.torch
        ptr, cap = self.ctx.buffer_ptr(name)
        nbytes = int(count) * np.dtype(dtype).itemsize
        out = np.empty(int(count), dtype=dtype)
        if nbytes == 0:
            return out
        _hip_memcpy_d2h(out.ctypes.data, ptr, nbytes)
        return out


class Batch:
    """Many frames through fsr_batch_run: per-frame device buffers (pristine
    input, working copy, field, soup, stats) and one descriptor each."""

    def __init__(self, method: str, frames, lanes: int = 8, device: int = 0,
                 aniso: dict | None = None, ctx: Context | None = None, cap_factor: float = 4.0):
        import torch

        self.torch = torch
        self.rec = Reconstructor(method, device=device, ctx=ctx, aniso=aniso)
        self.ctx = self.rec.ctx
        self.method = method
        self.frames = frames
        self.dev = self.rec.device
        h = ctypes.c_void_p()
        _lib.check(self.ctx.lib.fsr_batch_create(self.ctx.handle, lanes, ctypes.byref(h)))
        self.handle = h
        self.bufs = []
        for fr in frames:
            cap = int(max(3 * 1024, cap_factor * fr.n))
            b = dict(src_pos=torch.from_numpy(fr.pos).to(self.dev),
                     src_dens=torch.from_numpy(fr.dens).to(self.dev))
            b["pos"] = torch.empty_like(b["src_pos"])
            b["dens"] = torch.empty_like(b["src_dens"])
            b["phi"] = torch.empty(fr.sc.count, dtype=torch.float32, device=self.dev)
            b["tri"] = torch.empty((cap, 3), dtype=torch.float32, device=self.dev)
            b["nrm"] = torch.empty((cap, 3), dtype=torch.float32, device=self.dev)
            b["stats"] = torch.zeros(_lib.STATS_WORDS, dtype=torch.int32, device=self.dev)
            b["cap"] = cap
            self.bufs.append(b)
        self.ts = torch.zeros(2 * len(frames), dtype=torch.int64, device=self.dev)
        self._build_descs()

    def _build_descs(self):
        n = len(self.frames)
        self.descs = (_lib.FrameDesc * n)()
        for i, (fr, b) in enumerate(zip(self.frames, self.bufs)):
            d = self.descs[i]
            d.src_pos, d.src_dens = b["src_pos"].data_ptr(), b["src_dens"].data_ptr()
            d.pos, d.dens = b["pos"].data_ptr(), b["dens"].data_ptr()
            d.num_particles = fr.n
            d.tri_capacity = b["cap"]
            d.sp_grid, d.sc_grid = grid_info(fr.sp), grid_info(fr.sc)
            d.sim = self.rec.params(fr)
            d.phi = b["phi"].data_ptr()
            d.tri_pos, d.tri_nrm = b["tri"].data_ptr(), b["nrm"].data_ptr()
            d.stats = b["stats"].data_ptr()

    def run(self, stream=None, timestamps: bool = True):
        torch = self.torch
        s = stream if stream is not None else torch.cuda.current_stream(self.dev)
        ap = self.rec.aniso_param()
        _lib.check(self.ctx.lib.fsr_batch_run(
            self.handle, METHODS[self.method], ctypes.byref(ap), self.descs, len(self.frames),
            ctypes.c_void_p(self.ts.data_ptr() if timestamps else 0),
            ctypes.c_void_p(s.cuda_stream)))

    def grow_overflowed(self) -> bool:
        """After a synchronised run: grows the soup of frames that overflowed;
        returns True when descriptors changed (the next run recaptures)."""
        changed = False
        for b in self.bufs:
            st = stats_dict(b["stats"])
            if st["overflow"]:
                cap = st["num_mesh_vertices"] + 1024
                b["tri"] = self.torch.empty((cap, 3), dtype=self.torch.float32, device=self.dev)
                b["nrm"] = self.torch.empty((cap, 3), dtype=self.torch.float32, device=self.dev)
                b["cap"] = cap
                changed = True
        if changed:
            self._build_descs()
        return changed

    def frame_ms(self) -> np.ndarray:
        """Per-frame device time (start/end timestamps, 100 MHz clock)."""
        t = self.ts.cpu().numpy().reshape(-1, 2)
        return (t[:, 1] - t[:, 0]) * 1e-5

    def close(self):
        if self.handle:
            self.ctx.lib.fsr_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()


_hip = None


def _hip_memcpy_d2h(dst: int, src: int, nbytes: int) -> None:
    global _hip
    if _hip is None:
        import torch  # noqa: F401  (loads the HIP runtime torch uses)

        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipDeviceSynchronize.restype = ctypes.c_int
    _hip.hipDeviceSynchronize()
    rc = _hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 2)  # DeviceToHost
    if rc != 0:
        raise RuntimeError("hipMemcpy D2H failed with %d" % rc)
