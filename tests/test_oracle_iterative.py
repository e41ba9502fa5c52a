"""Pins the iterative-loop oracle (python/SART.py, python/ASD_POCS.py restatements)."""
import numpy as np

from oracle import dd_oracle as O
from oracle import iterative_oracle as I
from tests.conftest import make_params


def tv_functional(x, eps):
    # sum over voxels with all backward neighbours of sqrt(eps + |backward gradient|^2)
    g = x[1:, 1:, 1:]
    d0 = g - x[:-1, 1:, 1:]
    d1 = g - x[1:, :-1, 1:]
    d2 = g - x[1:, 1:, :-1]
    return np.sum(np.sqrt(eps + d0 ** 2 + d1 ** 2 + d2 ** 2))


def test_tv_grad_is_normalised_gradient_of_tv_functional():
    rng = np.random.default_rng(0)
    x = rng.random((6, 7, 5))
    eps = 1e-7
    h = 1e-6
    num = np.zeros_like(x)
    for idx in np.ndindex(x.shape):
        e = np.zeros_like(x)
        e[idx] = h
        num[idx] = (tv_functional(x + e, eps) - tv_functional(x - e, eps)) / (2 * h)
    inner = num[1:-1, 1:-1, 1:-1]
    expect = np.pad(inner / np.sqrt(np.sum(inner ** 2)), 1)
    np.testing.assert_allclose(I.tv_grad(x, eps), expect, rtol=0, atol=1e-6)


def test_sart_single_voxel_update_by_hand():
    # one voxel, one view, one detector cell: FP x = L x w, BP r = dtheta*W*r (W the FDK weight)
    p = make_params(NumberOfImage=[1, 1, 1], ImagePixelSpacing=[2.0, 2.0, 2.0], NumberOfDetectorPixels=[1, 1],
                    DetectorPixelSize=[4.0, 4.0], NumberOfViews=1, EndAngle=2 * np.pi)
    g = O.Geometry(p)
    a = O.forward_project(g, np.ones((1, 1, 1)))[0, 0, 0]      # A 1
    b = O.backproject(g, np.ones((1, 1, 1)), "ref")[0, 0, 0]   # B 1
    p0 = np.array([[[3.0]]])
    x1 = I.sart_reference(g, p0, 1, alpha=1.0, eps=1e-5)[0, 0, 0]
    # x1 = 0 - (B((0 - p)/(a + eps)) / b + eps) = p/(a+eps) - eps  (B is linear, B r = b r)
    assert np.isclose(x1, 3.0 / (a + 1e-5) - 1e-5, rtol=1e-12)


def test_sart_fixed_point_on_consistent_data():
    p = make_params(NumberOfImage=[6, 6, 4], NumberOfDetectorPixels=[12, 8], NumberOfViews=8)
    g = O.Geometry(p)
    x_true = np.full((4, 6, 6), 0.7)
    proj = O.forward_project(g, x_true)
    x = I.sart_reference(g, proj, 30, alpha=1.0, eps=1e-5)
    assert np.sqrt(np.mean((x - x_true) ** 2)) < 0.05
