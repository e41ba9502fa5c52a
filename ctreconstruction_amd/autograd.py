"""The DD projector as a differentiable torch operator.

``DDProjectorFunction`` computes y = A x with the HIP forward projector; its backward is the
hand-written matched backprojector A^T (``ctr_dd_backproject`` in adjoint mode), so gradients
of any loss of A x flow through the project's own kernels.  ``DDBackprojectorFunction`` is the
transpose pair (y -> A^T y, backward A).  Both accept one volume [nz, ny, nx] or a batch
[B, nz, ny, nx] (projections [nView, nv, nu] or [B, nView, nv, nu]); a batch is processed one
item at a time on the current stream.
"""
from __future__ import annotations

import torch

from . import projector as P
from .geometry import ScanGeometry


def _batched(fn, x: torch.Tensor, item_shape, out_shape, geom):
    x = x.contiguous()
    if tuple(x.shape) == tuple(item_shape):
        return fn(x, geom)
    if x.dim() == len(item_shape) + 1 and tuple(x.shape[1:]) == tuple(item_shape):
        out = torch.empty((x.shape[0],) + tuple(out_shape), dtype=torch.float32, device=x.device)
        for b in range(x.shape[0]):
            fn(x[b], geom, out=out[b])
        return out
    raise ValueError(f"expected shape {tuple(item_shape)} or [B, *{tuple(item_shape)}], got {tuple(x.shape)}")


def forward_batch(vol: torch.Tensor, geom: ScanGeometry) -> torch.Tensor:
    return _batched(P.dd_forward, vol, geom.vol_shape, geom.proj_shape, geom)


def adjoint_batch(proj: torch.Tensor, geom: ScanGeometry) -> torch.Tensor:
    return _batched(P.dd_adjoint, proj, geom.proj_shape, geom.vol_shape, geom)


class DDProjectorFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol, geom):
        ctx.geom = geom
        return forward_batch(vol.detach().to(torch.float32), geom)

    @staticmethod
    def backward(ctx, grad):
        return adjoint_batch(grad.to(torch.float32), ctx.geom), None


class DDBackprojectorFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, proj, geom):
        ctx.geom = geom
        return adjoint_batch(proj.detach().to(torch.float32), geom)

    @staticmethod
    def backward(ctx, grad):
        return forward_batch(grad.to(torch.float32), ctx.geom), None


class DDProjector(torch.nn.Module):
    """``y = A x`` as a module; ``.T(y)`` is the matched backprojector ``A^T y``."""

    def __init__(self, geom: ScanGeometry | dict):
        super().__init__()
        self.geom = geom if isinstance(geom, ScanGeometry) else ScanGeometry(geom)
        self.geom.check_supported()

    def forward(self, vol: torch.Tensor) -> torch.Tensor:
        return DDProjectorFunction.apply(vol, self.geom)

    def T(self, proj: torch.Tensor) -> torch.Tensor:
        return DDBackprojectorFunction.apply(proj, self.geom)
