"""GPU tests of the bfloat16 MFMA convolution kernels and of the learned prior network.

Convolutions are compared with float32 CPU convolutions of the same bfloat16-rounded operands
(this also checks the MFMA operand/result lane layout); the network's loss and gradients are
compared with the CPU oracle, which rounds to bfloat16 at the same points.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from ctreconstruction_amd import _native as N
from ctreconstruction_amd import prior as PR
from oracle import prior_oracle as PO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    N.lib()
    return torch.device("cuda:0")


def ref_conv(x_bf, w_bf, b=None):
    y = F.conv2d(x_bf.float().cpu().permute(0, 3, 1, 2), w_bf.float().cpu().permute(0, 3, 1, 2),
                 None if b is None else b.float().cpu(), padding=1)
    return y.permute(0, 2, 3, 1)


@pytest.mark.parametrize("n,h,w,cin,cout", [(2, 17, 23, 64, 64), (1, 32, 32, 3, 64), (3, 9, 40, 64, 1),
                                            (1, 5, 7, 1, 64), (2, 12, 12, 64, 16), (1, 64, 64, 32, 48)])
def test_conv_forward_matches_cpu(dev, n, h, w, cin, cout):
    g = torch.Generator().manual_seed(n * 1000 + cin + cout)
    x = torch.randn((n, h, w, cin), generator=g).to(torch.bfloat16)
    wt = (torch.randn((cout, 3, 3, cin), generator=g) * 0.1).to(torch.bfloat16)
    b = torch.randn(cout, generator=g)
    out = PR.conv3x3_forward(x.to(dev), wt.to(dev), b.to(dev)).float().cpu()
    ref = ref_conv(x, wt, b)
    err = (out - ref).abs().max().item() / ref.abs().max().item()
    print(f"conv fwd {n}x{h}x{w} {cin}->{cout}: max rel err {err:.2e}")
    assert err <= 2 ** -7  # bfloat16 output rounding


@pytest.mark.parametrize("cin,cout,relu,resid", [(64, 64, True, True), (3, 64, True, False),
                                                 (1, 64, False, True), (64, 16, True, False),
                                                 (4, 8, True, True), (64, 1, False, False)])
def test_conv_forward_fused_epilogue(dev, cin, cout, relu, resid):
    """Bias, ReLU and residual in each kernel's epilogue: the persistent kernel (64 -> 64 and
    64 -> at most 16), the small-input kernel (at most 4 channels in) and the generic one."""
    g = torch.Generator().manual_seed(7 + cin + cout)
    x = torch.randn((2, 19, 37, cin), generator=g).to(torch.bfloat16)
    wt = (torch.randn((cout, 3, 3, cin), generator=g) * (0.3 / cin ** 0.5)).to(torch.bfloat16)
    b = torch.randn(cout, generator=g)
    r = torch.randn((2, 19, 37, cout), generator=g).to(torch.bfloat16) if resid else None
    out = PR.conv3x3_forward(x.to(dev), wt.to(dev), b.to(dev), relu=relu,
                             residual=r.to(dev) if resid else None).float().cpu()
    ref = ref_conv(x, wt, b)
    if relu:
        ref = torch.relu(ref)
    if resid:
        ref = ref + r.float()
    assert (out - ref).abs().max().item() <= 2 ** -7 * ref.abs().max().item()


@pytest.mark.parametrize("cin,cout", [(64, 64), (3, 64), (64, 1)])
def test_conv_weight_and_input_gradients_match_cpu_autograd(dev, cin, cout):
    g = torch.Generator().manual_seed(cin * 7 + cout)
    x = torch.randn((2, 15, 18, cin), generator=g).to(torch.bfloat16)
    wt = (torch.randn((cout, 3, 3, cin), generator=g) * 0.1).to(torch.bfloat16)
    gy = torch.randn((2, 15, 18, cout), generator=g).to(torch.bfloat16)
    gw, gb = PR.conv3x3_wgrad(x.to(dev), gy.to(dev))
    gx = PR.conv3x3_forward(gy.to(dev), PR.flip_transpose(wt.to(dev))).float().cpu()
    xr = x.float().requires_grad_(True)
    wr = wt.float().requires_grad_(True)
    br = torch.zeros(cout, requires_grad=True)
    y = F.conv2d(xr.permute(0, 3, 1, 2), wr.permute(0, 3, 1, 2), br, padding=1).permute(0, 2, 3, 1)
    y.backward(gy.float())
    ew = (gw.cpu() - wr.grad).abs().max().item() / wr.grad.abs().max().item()
    eb = (gb.cpu() - br.grad).abs().max().item() / br.grad.abs().max().item()
    ex = (gx - xr.grad).abs().max().item() / xr.grad.abs().max().item()
    print(f"wgrad {cin}->{cout}: weight {ew:.2e} bias {eb:.2e} input {ex:.2e}")
    assert ew <= 1e-5 and eb <= 1e-5
    assert ex <= 2 ** -7


def _oracle_params(net):
    cpu = lambda p: p.detach().float().cpu().clone().requires_grad_(True)  # noqa: E731
    return {"w_in": cpu(net.w_in), "b_in": cpu(net.b_in), "wa": [cpu(p) for p in net.wa],
            "ba": [cpu(p) for p in net.ba], "wb": [cpu(p) for p in net.wb], "bb": [cpu(p) for p in net.bb],
            "w_out": cpu(net.w_out), "b_out": cpu(net.b_out)}


def test_network_forward_and_gradients_match_cpu_oracle(dev):
    net = PR.ResidualPrior(seed=3).to(dev)
    assert 550_000 < net.n_params() < 650_000
    g = torch.Generator().manual_seed(11)
    x = torch.randn((2, 24, 20, 3), generator=g)
    target = torch.randn((2, 24, 20), generator=g)
    out = net(x.to(dev))
    loss = torch.mean((out - target.to(dev)) ** 2)
    loss.backward()
    prm = _oracle_params(net)
    out_ref = PO.forward(prm, x)
    loss_ref = torch.mean((out_ref - target) ** 2)
    loss_ref.backward()
    e_out = (out.detach().cpu() - out_ref.detach()).abs().max().item() / out_ref.abs().max().item()
    assert e_out <= 2e-2, e_out
    assert abs(loss.item() - loss_ref.item()) <= 2e-2 * loss_ref.item()
    # gradients: cosine similarity and relative norm per parameter tensor (bfloat16 backprop)
    pairs = [(net.w_in, prm["w_in"]), (net.w_out, prm["w_out"]), (net.b_out, prm["b_out"])]
    pairs += [(a, b) for a, b in zip(net.wa, prm["wa"])] + [(a, b) for a, b in zip(net.wb, prm["wb"])]
    worst = 1.0
    for p_gpu, p_ref in pairs:
        a = p_gpu.grad.detach().cpu().flatten().double()
        b = p_ref.grad.detach().flatten().double()
        cos = float(torch.dot(a, b) / (a.norm() * b.norm()))
        worst = min(worst, cos)
        assert cos >= 0.995 and abs(a.norm() / b.norm() - 1) <= 0.03
    print(f"network: out rel err {e_out:.2e}, worst gradient cosine {worst:.5f}")


def test_fused_inference_matches_training_forward(dev):
    net = PR.ResidualPrior(seed=5).to(dev)
    x = torch.randn((3, 33, 31, 3), device=dev)
    with torch.no_grad():
        a = net(x)
    b = net.infer(x)
    assert (a - b).abs().max().item() <= 2e-2 * a.abs().max().item()
