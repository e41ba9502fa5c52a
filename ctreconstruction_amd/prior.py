"""Learned 2.5D residual prior: a small CNN in bfloat16 on MFMA with FP32 accumulation.

No counterpart in the reference.  The network maps three adjacent slices of a sparse-view
reconstruction to a correction of the centre slice:

    h  = relu(conv(x, W_in) + b_in)                       3 -> 64 channels
    h  = h + conv(relu(conv(h, W_a) + b_a), W_b) + b_b     8 residual blocks, 64 -> 64
    y  = x_centre + conv(h, W_out) + b_out                 64 -> 1

(3x3 convolutions, ~0.59 M parameters).  Activations are NHWC bfloat16 [N, H, W, C]; master
weights are FP32 (updated by the optimiser) and cast to bfloat16 for the kernels.  Every
convolution, its input gradient and its weight gradient run in the hand-written kernels of
``csrc/prior.hip`` (``ctr_conv3x3_bf16_*``); ReLU and the residual adds are torch elementwise
ops.  ``infer`` runs the same network with bias, ReLU and residual fused into the conv epilogue.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _native as N

BF16 = torch.bfloat16


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def conv3x3_forward(x: torch.Tensor, w_bf16: torch.Tensor, bias: torch.Tensor | None = None,
                    relu: bool = False, residual: torch.Tensor | None = None) -> torch.Tensor:
    """y[N,H,W,Cout] = act(conv3x3(x) + bias) + residual, bfloat16 in/out, FP32 accumulation."""
    if x.dtype != BF16 or w_bf16.dtype != BF16 or x.dim() != 4 or not x.is_contiguous():
        raise TypeError("x must be contiguous NHWC bfloat16 and weights bfloat16")
    n, h, w, cin = x.shape
    cout = w_bf16.shape[0]
    if tuple(w_bf16.shape) != (cout, 3, 3, cin):
        raise ValueError(f"weight shape {tuple(w_bf16.shape)} != ({cout}, 3, 3, {cin})")
    out = torch.empty((n, h, w, cout), dtype=BF16, device=x.device)
    if residual is not None:
        residual = residual.contiguous()
        if residual.dtype != BF16 or tuple(residual.shape) != tuple(out.shape):
            raise ValueError("residual must be bfloat16 with the output's shape")
    b = bias.contiguous().float() if bias is not None else None
    N.check(N.lib().ctr_conv3x3_bf16_forward(_ptr(x), _ptr(w_bf16.contiguous()), _ptr(b), _ptr(residual),
                                              _ptr(out), n, h, w, cin, cout, int(relu), _stream(x.device)),
            "ctr_conv3x3_bf16_forward")
    return out


def conv3x3_wgrad(x: torch.Tensor, gout: torch.Tensor, with_bias: bool = True):
    """(grad_weight [Cout,3,3,Cin] float32, grad_bias [Cout] float32) summed over all pixels."""
    n, h, w, cin = x.shape
    cout = gout.shape[-1]
    lib = N.lib()
    gw = torch.empty((cout, 3, 3, cin), dtype=torch.float32, device=x.device)
    gb = torch.empty(cout, dtype=torch.float32, device=x.device) if with_bias else None
    nbytes = lib.ctr_conv3x3_bf16_wgrad_workspace_bytes(n, h, w, cin, cout)
    ws = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=x.device)
    N.check(lib.ctr_conv3x3_bf16_wgrad(_ptr(x), _ptr(gout.contiguous()), _ptr(gw), _ptr(gb), n, h, w, cin, cout,
                                        _ptr(ws), nbytes, _stream(x.device)), "ctr_conv3x3_bf16_wgrad")
    return gw, gb


def flip_transpose(w_bf16: torch.Tensor) -> torch.Tensor:
    """Weights of the input-gradient convolution: w'[ci, ky, kx, co] = w[co, 2-ky, 2-kx, ci]."""
    return torch.flip(w_bf16, dims=(1, 2)).permute(3, 1, 2, 0).contiguous()


class Conv3x3Function(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, weight, bias, need_input_grad):
        wb = weight.detach().to(BF16).contiguous()
        ctx.save_for_backward(x, wb)
        ctx.need_input_grad = need_input_grad
        return conv3x3_forward(x, wb, bias.detach())

    @staticmethod
    def backward(ctx, gout):
        x, wb = ctx.saved_tensors
        gout = gout.to(BF16).contiguous()
        gx = conv3x3_forward(gout, flip_transpose(wb)) if ctx.need_input_grad else None
        gw, gb = conv3x3_wgrad(x, gout)
        return gx, gw, gb, None


class ResidualPrior(torch.nn.Module):
    """8-block 2.5D residual CNN (3 adjacent slices in, 64 channels, 3x3 convolutions)."""

    def __init__(self, channels: int = 64, blocks: int = 8, in_slices: int = 3, seed: int = 0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)

        def conv(cin, cout, scale=1.0):
            std = scale * math.sqrt(2.0 / (9 * cin))
            w = torch.randn((cout, 3, 3, cin), generator=g) * std
            return torch.nn.Parameter(w), torch.nn.Parameter(torch.zeros(cout))

        self.in_slices = in_slices
        self.w_in, self.b_in = conv(in_slices, channels)
        self.wa, self.ba, self.wb, self.bb = (torch.nn.ParameterList() for _ in range(4))
        for _ in range(blocks):
            w, b = conv(channels, channels)
            self.wa.append(w)
            self.ba.append(b)
            w, b = conv(channels, channels, scale=0.1)  # residual branches start near identity
            self.wb.append(w)
            self.bb.append(b)
        self.w_out, self.b_out = conv(channels, 1, scale=0.1)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """x: [N, H, W, 3] (any float dtype) -> corrected centre slice [N, H, W] float32."""
        xb = x.to(BF16).contiguous()
        h = torch.relu(Conv3x3Function.apply(xb, self.w_in, self.b_in, False))
        for wa, ba, wb, bb in zip(self.wa, self.ba, self.wb, self.bb):
            t = torch.relu(Conv3x3Function.apply(h, wa, ba, True))
            h = h + Conv3x3Function.apply(t, wb, bb, True)
        y = Conv3x3Function.apply(h, self.w_out, self.b_out, True)
        c = self.in_slices // 2
        return x[..., c].float() + y[..., 0].float()

    @torch.no_grad()
    def infer(self, x: torch.Tensor) -> torch.Tensor:
        """Inference with bias, ReLU and residual fused into the convolution epilogues."""
        xb = x.to(BF16).contiguous()
        bf = lambda t: t.detach().to(BF16).contiguous()  # noqa: E731
        h = conv3x3_forward(xb, bf(self.w_in), self.b_in.detach(), relu=True)
        for wa, ba, wb, bb in zip(self.wa, self.ba, self.wb, self.bb):
            t = conv3x3_forward(h, bf(wa), ba.detach(), relu=True)
            h = conv3x3_forward(t, bf(wb), bb.detach(), residual=h)
        y = conv3x3_forward(h, bf(self.w_out), self.b_out.detach())
        c = self.in_slices // 2
        return x[..., c].float() + y[..., 0].float()

    def n_params(self) -> int:
        return sum(p.numel() for p in self.parameters())


def slices_to_input(vol: torch.Tensor, z: torch.Tensor | None = None, in_slices: int = 3) -> torch.Tensor:
    """2.5D input [len(z), ny, nx, in_slices] of adjacent slices (clamped at the volume ends)."""
    nz = vol.shape[0]
    if z is None:
        z = torch.arange(nz, device=vol.device)
    half = in_slices // 2
    idx = torch.stack([torch.clamp(z + d, 0, nz - 1) for d in range(-half, half + 1)], dim=-1)
    return vol[idx].permute(0, 2, 3, 1).contiguous()


@torch.no_grad()
def apply_to_volume(net: ResidualPrior, vol: torch.Tensor, scale: float = 1.0, batch: int = 256) -> torch.Tensor:
    """Run the prior slice by slice over a volume [nz, ny, nx]; inputs are multiplied by ``scale``
    (network units) and the output divided by it.  ``batch`` slices go through each layer call
    (256 slices of 256^2 hold ~2 GB per 64-channel activation)."""
    out = torch.empty_like(vol)
    nz = vol.shape[0]
    for z0 in range(0, nz, batch):
        z = torch.arange(z0, min(z0 + batch, nz), device=vol.device)
        out[z0:z0 + len(z)] = net.infer(slices_to_input(vol, z, net.in_slices) * scale) / scale
    return out
