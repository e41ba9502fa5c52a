"""GPU tests of the matched backprojector, the autograd projector and the iterative loops.

* A^T against the oracle's exact transpose (FP64 scatter of the same weights), several geometries;
* the adjoint identity <Ax, y> = <x, A^T y> and a central finite-difference check of
  ||Ax - b||^2 at the differentiable-projector configuration (8 x 128^3, 180 views, 256^2);
* SART / ASD-POCS: reference loop against the NumPy oracle, TV gradient against the oracle and
  against autograd, and convergence of the autograd-driven loops.
"""
import numpy as np
import pytest
import torch

import ctreconstruction_amd as ctr
from ctreconstruction_amd import _native as N
from ctreconstruction_amd import iterative as IT
from ctreconstruction_amd import phantom
from ctreconstruction_amd.autograd import DDProjector, DDProjectorFunction
from oracle import dd_oracle as O
from oracle import iterative_oracle as IO
from tests.conftest import make_params
from tests.test_gpu_parity import CASES

pytestmark = pytest.mark.gpu

C4 = dict(NumberOfImage=[128, 128, 128], ImagePixelSpacing=[0.5, 0.5, 0.5], NumberOfDetectorPixels=[256, 256],
          DetectorPixelSize=[0.75, 0.75], NumberOfViews=180)


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    N.lib()
    return torch.device("cuda:0")


def _rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("name", ["flat", "flat_45deg_ties", "flat_offsets", "anisotropic", "helical",
                                  "curved", "wide_footprint", "curved_fan_2d", "near_source", "off_center",
                                  "reverse", "helical_reverse", "short_scan", "source_z", "origin",
                                  "single_view", "thin_z", "curved_helical", "curved_offsets", "curved_reverse",
                                  "curved_near_source", "aniso_helical", "flat_fan_2d"])
def test_adjoint_matches_oracle_transpose(dev, name):
    p = make_params(**CASES[name])
    g = O.Geometry(p)
    y = np.random.default_rng(21).standard_normal((g.nview, g.nv, g.nu)).astype(np.float32)
    out = ctr.projector.dd_adjoint(torch.from_numpy(y).to(dev), ctr.ScanGeometry(p)).cpu().numpy()
    ref = O.forward_adjoint(g, y.astype(np.float64))
    err = _rel(out, ref)
    print(f"adjoint[{name}] max rel err {err:.2e}")
    assert err <= 1e-4


def _dot_pair(A, x, y):
    with torch.no_grad():
        ax = A(x)
        aty = A.T(y)
    dims = tuple(range(1, x.dim()))
    pdims = tuple(range(1, y.dim()))
    lhs = torch.sum(ax.double() * y.double(), dim=pdims)
    rhs = torch.sum(x.double() * aty.double(), dim=dims)
    scale = torch.sum(ax.double().abs() * y.double().abs(), dim=pdims)
    return lhs, rhs, scale


def test_adjoint_identity_at_batch_configuration(dev):
    """<Ax, y> = <x, A^T y> per batch item.  With Gaussian x, y the inner products cancel, so the
    difference is measured against sum |Ax| |y|, the scale of their rounding; with positive x, y
    nothing cancels and a weight mismatch anywhere adds up coherently."""
    geom = ctr.ScanGeometry(make_params(**C4))
    A = DDProjector(geom)
    gen = torch.Generator(device=dev)
    gen.manual_seed(404)
    x = torch.randn((8,) + geom.vol_shape, device=dev, generator=gen)
    y = torch.randn((8,) + geom.proj_shape, device=dev, generator=gen)
    lhs, rhs, scale = _dot_pair(A, x, y)
    err = ((lhs - rhs).abs() / scale).max().item()
    print(f"adjoint identity (Gaussian): <Ax,y> {lhs.tolist()[:2]} <x,A^T y> {rhs.tolist()[:2]}, "
          f"max |diff| / sum|Ax||y| {err:.2e}")
    assert err <= 1e-6
    xp = torch.rand((8,) + geom.vol_shape, device=dev, generator=gen)
    yp = torch.rand((8,) + geom.proj_shape, device=dev, generator=gen)
    lhs, rhs, _ = _dot_pair(A, xp, yp)
    rel = ((lhs - rhs).abs() / lhs.abs()).max().item()
    print(f"adjoint identity (positive): max rel {rel:.2e}")
    assert rel <= 1e-6


@pytest.mark.parametrize("eps,tol_max,tol_median", [(0.25, 5e-3, 1e-4), (1e-3, 5e-2, 1e-3)])
def test_finite_difference_gradient_of_data_term(dev, eps, tol_max, tol_median):
    """f(x) = ||A x - b||^2 ; autograd gradient (backward = A^T) vs central differences on 64
    voxels of the 128^3 / 180-view / 256^2 configuration.  f is quadratic in x, so a central
    difference is exact at any step up to rounding: at eps = 0.25 it pins the gradient at
    5e-3 of its size.  At the configured eps = 1e-3 the FP32 rounding of the forward projector
    shows: a z-prefix breakpoint holds a line integral to half an ulp, and a changed voxel
    re-rounds every breakpoint above it, so f moves by rounding noise beside the 2 eps g signal
    (measured: median 1.5e-4 to 3.5e-4, worst of 64 voxels 7e-3 to 2.8e-2 of the gradient).  A
    wrong gradient would be off by O(1)."""
    geom = ctr.ScanGeometry(make_params(**C4))
    ell = phantom.random_ellipsoids(8, phantom.half_extent(geom), 401)
    x = (phantom.ellipsoid_volume(geom, ell, device=dev) * 0.5).contiguous()
    gen = torch.Generator(device=dev).manual_seed(403)
    b = ctr.dd_forward(x, geom) + 0.1 * torch.randn(geom.proj_shape, device=dev, generator=gen)
    x0 = (x + 0.05 * torch.randn(x.shape, device=dev, generator=gen)).contiguous()
    xv = x0.clone().requires_grad_(True)
    f = torch.sum((DDProjectorFunction.apply(xv, geom) - b).double() ** 2)
    (grad,) = torch.autograd.grad(f, xv)

    def fval(v):
        with torch.no_grad():
            return torch.sum((ctr.dd_forward(v.contiguous(), geom) - b).double() ** 2).item()

    rng = np.random.default_rng(402)
    idx = [tuple(int(rng.integers(16, n - 16)) for n in geom.vol_shape) for _ in range(64)]
    num, ana = [], []
    for i in idx:
        xp = x0.clone()
        xp[i] += eps
        xm = x0.clone()
        xm[i] -= eps
        num.append((fval(xp) - fval(xm)) / (2 * eps))
        ana.append(float(grad[i]))
    num, ana = np.array(num), np.array(ana)
    err = np.abs(num - ana) / np.maximum(np.abs(ana), 1e-3 * np.abs(ana).max())
    print(f"finite-difference check, eps {eps}: max rel err {err.max():.2e}, median {np.median(err):.2e}")
    assert err.max() <= tol_max and np.median(err) <= tol_median


def test_autograd_backward_is_matched_backprojector(dev):
    geom = ctr.ScanGeometry(make_params(**CASES["flat_offsets"]))
    x = torch.rand(geom.vol_shape, device=dev, requires_grad=True)
    g = torch.randn(geom.proj_shape, device=dev)
    y = DDProjectorFunction.apply(x, geom)
    y.backward(g)
    torch.testing.assert_close(x.grad, ctr.projector.dd_adjoint(g, geom), rtol=0, atol=0)
    p = torch.randn(geom.proj_shape, device=dev, requires_grad=True)
    A = DDProjector(geom)
    v = A.T(p)
    gv = torch.randn(geom.vol_shape, device=dev)
    v.backward(gv)
    torch.testing.assert_close(p.grad, ctr.dd_forward(gv, geom), rtol=0, atol=0)


def test_sart_reference_loop_matches_oracle(dev):
    p = make_params(NumberOfImage=[12, 12, 8], NumberOfDetectorPixels=[24, 16], NumberOfViews=10)
    g = O.Geometry(p)
    x_true = np.random.default_rng(22).random((8, 12, 12))
    proj = O.forward_project(g, x_true)
    ref = IO.sart_reference(g, proj, 3, alpha=1.0, eps=1e-5)
    geom = ctr.ScanGeometry(p)
    out = IT.sart_reference(torch.from_numpy(proj.astype(np.float32)).to(dev), geom, 3, alpha=1.0,
                            eps=1e-5).cpu().numpy()
    err = _rel(out, ref)
    print(f"sart_reference max rel err {err:.2e}")
    assert err <= 1e-4


def test_tv_gradient_matches_oracle_and_autograd(dev):
    x = np.random.default_rng(23).random((20, 18, 16)).astype(np.float32)
    xt = torch.from_numpy(x).to(dev)
    a = IT.tv_gradient(xt).cpu().numpy()
    b = IT.tv_gradient_autograd(xt).cpu().numpy()
    ref = IO.tv_grad(x.astype(np.float64))
    assert np.abs(a - ref).max() <= 1e-5 * np.abs(ref).max()
    assert np.abs(b - ref).max() <= 1e-5 * np.abs(ref).max()


def test_sart_50_iterations_through_autograd_on_batch(dev):
    geom = ctr.ScanGeometry(make_params(**C4))
    vols = torch.stack([phantom.ellipsoid_volume(geom, phantom.random_ellipsoids(
        8, phantom.half_extent(geom), 400 + b), device=dev) for b in range(8)])
    vols = torch.clamp_min(vols, 0.0)
    from ctreconstruction_amd.autograd import forward_batch
    proj = forward_batch(vols, geom)
    res = []

    def cb(it, x):
        if it in (0, 9, 49):
            r = forward_batch(x, geom) - proj
            res.append((float(torch.sqrt(torch.mean(r ** 2))), float(torch.sqrt(torch.mean((x - vols) ** 2)))))

    x = IT.sart(proj, geom, 50, relaxation=0.5, backprojector="matched", callback=cb)
    print("SART (residual rms, rmse):", res)
    assert res[2][0] < 0.2 * res[0][0]
    assert res[2][1] < res[0][1]
    assert torch.isfinite(x).all()


def _tv(x):
    d = torch.sqrt(1e-7 + (x[..., 1:, 1:, 1:] - x[..., :-1, 1:, 1:]) ** 2
                   + (x[..., 1:, 1:, 1:] - x[..., 1:, :-1, 1:]) ** 2
                   + (x[..., 1:, 1:, 1:] - x[..., 1:, 1:, :-1]) ** 2)
    return torch.sum(d, dim=(-3, -2, -1))


def test_asd_pocs_tv_steps_through_autograd(dev):
    """ASD-POCS (python/ASD_POCS.py:167-199) on the batch configuration: 8 volumes of 128^3,
    180 views, 256^2 cells; 20 TV steps per outer iteration, the TV gradient by autograd."""
    geom = ctr.ScanGeometry(make_params(**C4))
    vols = torch.stack([torch.clamp_min(phantom.ellipsoid_volume(geom, phantom.random_ellipsoids(
        8, phantom.half_extent(geom), 420 + b), device=dev), 0.0) for b in range(8)])
    from ctreconstruction_amd.autograd import forward_batch
    proj = forward_batch(vols, geom)
    kw = dict(na=5, backprojector="matched")
    # one outer iteration (FDK start, 5 POCS steps, 20 TV steps), autograd vs explicit TV gradient
    x_pocs = IT.asd_pocs(proj, geom, n_outer=1, ng=0, **kw)
    x_tv = IT.asd_pocs(proj, geom, n_outer=1, ng=20, tv="autograd", **kw)
    x_tv_ref = IT.asd_pocs(proj, geom, n_outer=1, ng=20, tv="reference", **kw)
    assert (x_tv - x_tv_ref).abs().max() <= 1e-3 * x_tv_ref.abs().max()
    # the normalised TV gradient is a descent direction of every volume's TV
    xd = x_pocs.double()
    g = IT.tv_gradient_autograd(xd)
    tv0, tv_step = _tv(xd), _tv(xd - 0.05 * g)
    print(f"ASD-POCS TV per volume: POCS {tv0.tolist()}, one 0.05 step {tv_step.tolist()}, "
          f"20 TV steps {_tv(x_tv.double()).tolist()}")
    assert (tv_step < tv0).all()
    # two outer iterations against plain non-negative SART
    x = IT.asd_pocs(proj, geom, n_outer=2, ng=20, tv="autograd", **kw)
    rmse = float(torch.sqrt(torch.mean((x - vols) ** 2)))
    x_sart = IT.sart(proj, geom, 10, relaxation=1.0, nonneg=True)
    rmse_sart = float(torch.sqrt(torch.mean((x_sart - vols) ** 2)))
    print(f"ASD-POCS rmse {rmse:.4f} (10-step SART {rmse_sart:.4f})")
    # the TV steps are not followed by a positivity projection (python/ASD_POCS.py:186-188)
    assert torch.isfinite(x).all() and x.min() >= -0.01 * x.max()
    assert rmse < 1.5 * rmse_sart
