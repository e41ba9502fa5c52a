"""CPU oracle of the learned residual prior (torch float32 on the CPU).

TEST INFRASTRUCTURE ONLY.  The same network as ``ctreconstruction_amd.prior.ResidualPrior``
written with ``torch.nn.functional.conv2d`` in float32 on the CPU, rounding to bfloat16 where
the GPU path stores bfloat16 (input, weights, every convolution output, every activation), so
that GPU-vs-oracle differences come from accumulation order and rare rounding-boundary flips
only.  Gradients come from torch's CPU autograd.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F


def bf(t: torch.Tensor) -> torch.Tensor:
    return t.to(torch.bfloat16).to(torch.float32)


class _RoundBF16(torch.autograd.Function):
    """Round to bfloat16 in the forward pass; in the backward pass round the gradient too
    (the GPU path carries bfloat16 gradients between layers)."""

    @staticmethod
    def forward(ctx, x):
        return bf(x)

    @staticmethod
    def backward(ctx, g):
        return bf(g)


def conv(x_nhwc: torch.Tensor, w: torch.Tensor, b: torch.Tensor | None) -> torch.Tensor:
    """3x3 same convolution, NHWC float32 in/out, weights [Cout, 3, 3, Cin]."""
    y = F.conv2d(x_nhwc.permute(0, 3, 1, 2), bf(w).permute(0, 3, 1, 2), None if b is None else b, padding=1)
    return y.permute(0, 2, 3, 1)


def forward(params: dict, x: torch.Tensor) -> torch.Tensor:
    """params: w_in, b_in, wa[i], ba[i], wb[i], bb[i], w_out, b_out (float32 CPU tensors)."""
    r = _RoundBF16.apply
    xb = r(x)
    h = r(torch.relu(r(conv(xb, params["w_in"], params["b_in"]))))
    for wa, ba, wb, bb in zip(params["wa"], params["ba"], params["wb"], params["bb"]):
        t = r(torch.relu(r(conv(h, wa, ba))))
        h = r(h + r(conv(t, wb, bb)))
    y = r(conv(h, params["w_out"], params["b_out"]))
    c = x.shape[-1] // 2
    return x[..., c] + y[..., 0]
