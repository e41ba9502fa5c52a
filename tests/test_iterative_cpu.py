"""CPU checks of the host-side parts of the iterative and learned layers (torch on the CPU)."""
import numpy as np
import torch

from ctreconstruction_amd import iterative as IT
from ctreconstruction_amd import prior as PR
from oracle import iterative_oracle as IO


def test_tv_gradient_matches_oracle_on_cpu():
    x = np.random.default_rng(31).random((12, 10, 9))
    ref = IO.tv_grad(x)
    got = IT.tv_gradient(torch.from_numpy(x)).numpy()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


def test_tv_gradient_autograd_equals_explicit_formula():
    x = torch.from_numpy(np.random.default_rng(32).random((2, 11, 9, 8)))
    np.testing.assert_allclose(IT.tv_gradient_autograd(x).numpy(), IT.tv_gradient(x).numpy(), rtol=0, atol=1e-12)


def test_tv_gradient_of_constant_volume_is_zero():
    x = torch.full((6, 6, 6), 3.0, dtype=torch.float64)
    assert torch.count_nonzero(IT.tv_gradient(x)) == 0


def test_slices_to_input_clamps_at_volume_ends():
    vol = torch.arange(5 * 2 * 3, dtype=torch.float32).reshape(5, 2, 3)
    inp = PR.slices_to_input(vol)
    assert inp.shape == (5, 2, 3, 3)
    assert torch.equal(inp[0, ..., 0], vol[0]) and torch.equal(inp[0, ..., 1], vol[0])
    assert torch.equal(inp[2, ..., 0], vol[1]) and torch.equal(inp[2, ..., 2], vol[3])
    assert torch.equal(inp[4, ..., 2], vol[4])


def test_prior_network_size():
    net = PR.ResidualPrior()
    # 3->64 input conv, 8 blocks of two 64->64 convs, 64->1 output conv (+ biases)
    expect = (9 * 3 * 64 + 64) + 16 * (9 * 64 * 64 + 64) + (9 * 64 + 1)
    assert net.n_params() == expect
    assert 0.55e6 < expect < 0.65e6


def test_flip_transpose_is_the_adjoint_of_the_convolution():
    # <conv(x, w), y> = <x, conv(y, w')> with w' = flip_transpose(w), checked with float64 CPU convs
    import torch.nn.functional as F
    g = torch.Generator().manual_seed(5)
    w = torch.randn((4, 3, 3, 5), generator=g, dtype=torch.float64)
    x = torch.randn((1, 7, 6, 5), generator=g, dtype=torch.float64)
    y = torch.randn((1, 7, 6, 4), generator=g, dtype=torch.float64)

    def conv(a, k):
        return F.conv2d(a.permute(0, 3, 1, 2), k.permute(0, 3, 1, 2), padding=1).permute(0, 2, 3, 1)

    wt = PR.flip_transpose(w)
    lhs = torch.sum(conv(x, w) * y)
    rhs = torch.sum(x * conv(y, wt))
    assert abs(float(lhs - rhs)) < 1e-10 * abs(float(lhs))


def test_training_checkpoint_round_trip(tmp_path):
    from ctreconstruction_amd import learned as L
    net = PR.ResidualPrior(seed=2)
    opt = torch.optim.Adam(net.parameters(), lr=1e-3)
    g = torch.Generator().manual_seed(0)
    for p in net.parameters():
        p.grad = torch.randn(p.shape, generator=g)
    opt.step()
    path = str(tmp_path / "ck.pt")
    L.save_checkpoint(net, opt, 17, path)
    net2, opt2, nxt = L.load_checkpoint(path, 1e-3, device="cpu")
    assert nxt == 17
    for a, b in zip(net.parameters(), net2.parameters()):
        assert torch.equal(a.detach(), b.detach())
    s1, s2 = opt.state_dict()["state"], opt2.state_dict()["state"]
    assert s1.keys() == s2.keys()
    for k in s1:
        assert torch.equal(s1[k]["exp_avg"], s2[k]["exp_avg"])
        assert torch.equal(s1[k]["exp_avg_sq"], s2[k]["exp_avg_sq"])
