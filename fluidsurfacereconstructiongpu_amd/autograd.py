"""torch.autograd binding of the adjoint: particle positions -> (scalar field,
triangle-soup vertex positions), differentiable with respect to the positions.

The forward is one frame of libfsr (ZB05 or Akinci12, both evaluate the ZB05
field of updateScalarFieldValueZB05, ReconstructionCUDA.cu:230-289) run with the
adjoint state kept; the backward calls fsr_mc_vertex_backward (vertex
interpolation of generateTriangles, ReconstructionCUDA.cu:383-463, topology held
fixed, normals constant) and fsr_field_zb05_backward.  The reference has no
backward pass.  Everything runs in libfsr.so; there is no fallback.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .datasets import Frame
from .reconstructor import Context, Reconstructor, stats_dict


class SurfaceField:
    """Differentiable reconstruction of one particle configuration layout (its
    grids and constants come from `fr`; the positions are the autograd input).

    Owns its context: the adjoint state of a forward is valid until the next
    forward of the same object, which `backward` checks."""

    def __init__(self, fr: Frame, method: str = "zb05", device: int = 0,
                 tri_capacity: int | None = None):
        if method not in ("zb05", "akinci12"):
            raise ValueError("the adjoint covers the ZB05 field (methods zb05 and akinci12)")
        self.fr = fr
        self.method = method
        self.ctx = Context(device)
        _lib.check(self.ctx.lib.fsr_set_adjoint(self.ctx.handle, 1))
        self.rec = Reconstructor(method, device=device, ctx=self.ctx)
        self.device = self.rec.device
        self.dens = torch.from_numpy(fr.dens).to(self.device)
        self.cap = tri_capacity or max(3 * 1024, 6 * fr.n)
        self.sim = self.rec.params(fr)
        self.generation = 0
        self.stats: dict = {}

    def forward_frame(self, pos: torch.Tensor):
        """Runs one frame on a copy of pos (N, 3) float32; returns (phi (V,),
        soup (K, 3)) as fresh tensors.  Synchronises once (soup size)."""
        if pos.dtype != torch.float32 or pos.dim() != 2 or pos.shape[1] != 3:
            raise ValueError("pos must be a float32 (N, 3) tensor")
        if pos.shape[0] != self.fr.n:
            raise ValueError("pos has %d particles, the layout %d" % (pos.shape[0], self.fr.n))
        work = pos.detach().contiguous().clone()
        dens = self.dens.clone()
        self.rec.launch(self.fr, work, dens, self.cap, sim=self.sim)
        torch.cuda.synchronize(self.device)
        st = stats_dict(self.rec._stats)
        if st["overflow"]:
            self.cap = st["num_mesh_vertices"] + 1024
            work.copy_(pos.detach())
            dens.copy_(self.dens)
            self.rec.launch(self.fr, work, dens, self.cap, sim=self.sim)
            torch.cuda.synchronize(self.device)
            st = stats_dict(self.rec._stats)
        self.stats = st
        self.generation += 1
        V = self.fr.sc.count
        K = st["num_mesh_vertices"]
        phi = self.rec._phi[:V].clone()
        soup = self.rec._tri[:K].clone()
        return phi, soup

    def backward_frame(self, generation: int, g_phi, g_soup):
        if generation != self.generation:
            raise RuntimeError("SurfaceField: another forward ran since this graph was built; "
                               "the adjoint state of that frame is gone")
        V = self.fr.sc.count
        s = torch.cuda.current_stream(self.device)
        if g_phi is None:
            g = torch.zeros(V, dtype=torch.float32, device=self.device)
        else:
            g = g_phi.detach().to(torch.float32).contiguous().clone()
        if g_soup is not None and self.stats["num_mesh_vertices"] > 0:
            gs = g_soup.detach().to(torch.float32).contiguous()
            _lib.check(self.ctx.lib.fsr_mc_vertex_backward(
                self.ctx.handle, ctypes.c_void_p(gs.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                ctypes.c_void_p(s.cuda_stream)))
        out = torch.empty((self.fr.n, 3), dtype=torch.float32, device=self.device)
        _lib.check(self.ctx.lib.fsr_field_zb05_backward(
            self.ctx.handle, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(out.data_ptr()),
            ctypes.c_void_p(s.cuda_stream)))
        return out

    def __call__(self, pos: torch.Tensor):
        return _SurfaceFieldFn.apply(pos, self)

    def close(self):
        self.ctx.close()


class _SurfaceFieldFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, sf: SurfaceField):
        phi, soup = sf.forward_frame(pos)
        ctx.sf = sf
        ctx.generation = sf.generation
        ctx.mark_non_differentiable()
        return phi, soup

    @staticmethod
    def backward(ctx, g_phi, g_soup):
        return ctx.sf.backward_frame(ctx.generation, g_phi, g_soup), None


def sphere_loss(phi, soup, fr: Frame, radius: float = 0.3, w_grid: float = 1.0,
                w_soup: float = 1.0):
    """mean over the grid of (phi - SDF)^2 + mean over soup vertices of
    (|y| - radius)^2, SDF of the sphere of `radius` at the origin, evaluated at
    the grid's FP32 vertex positions."""
    dev = phi.device
    g = fr.sc
    ix = torch.arange(g.res[0], device=dev, dtype=torch.float32)
    iy = torch.arange(g.res[1], device=dev, dtype=torch.float32)
    iz = torch.arange(g.res[2], device=dev, dtype=torch.float32)
    cell = torch.tensor(float(g.cell), dtype=torch.float32, device=dev)
    x = torch.addcmul(torch.tensor(float(g.min[0]), device=dev), ix, cell).double()
    y = torch.addcmul(torch.tensor(float(g.min[1]), device=dev), iy, cell).double()
    z = torch.addcmul(torch.tensor(float(g.min[2]), device=dev), iz, cell).double()
    r = torch.sqrt(z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2)
    sdf = (r - radius).reshape(-1)
    loss = w_grid * torch.mean((phi.double() - sdf) ** 2)
    if soup.shape[0] and w_soup:
        loss = loss + w_soup * torch.mean((torch.linalg.norm(soup.double(), dim=1) - radius) ** 2)
    return loss
