"""Pins the pixel-driven FDK oracle (python/ConeBeamGPU.py:203-384) without reusing its code path."""
import math

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from oracle import pixel_oracle as P
from tests import analytic
from tests.conftest import make_cone_params
from tests.test_oracle_filter import closed_form_spectrum


@pytest.mark.parametrize("zp", [8, 64, 1024])
@pytest.mark.parametrize("tau", [0.26, 0.8])
@pytest.mark.parametrize("cutoff", [0.3, 0.5, 1.0])
def test_spectrum_is_twice_the_taps_cosine_series_with_sample_mask(zp, tau, cutoff):
    # ConeBeam.Filter: 2 |DFT(h)| and the mask |w| > pi*cutoff in rad/sample (no 1/(2 tau)).
    H = P.spectrum(zp + 1, tau, "ram-lak", cutoff)
    Hc, w = closed_form_spectrum(zp, tau)
    Hc = 2.0 * Hc
    Hc[np.abs(w) > np.pi * cutoff] = 0
    np.testing.assert_allclose(H, Hc, rtol=0, atol=1e-12 * Hc.max())
    assert H[np.abs(w) > np.pi * cutoff].max(initial=0.0) == 0.0


@pytest.mark.parametrize("ftype", P.FILTER_TYPES)
def test_windows_match_their_formulas_at_quarter_band(ftype):
    zp, tau, cutoff = 64, 0.5, 0.4
    H = P.spectrum(zp + 1, tau, ftype, cutoff)
    base = P.spectrum(zp + 1, tau, "ram-lak", cutoff)
    m = zp // 2 + zp // 8  # w = pi/4, inside the pass band pi*0.4
    w = np.pi / 4
    factor = {"ram-lak": 1.0, "shepp-logan": math.sin(w / (2 * cutoff)) / (w / (2 * cutoff)),
              "cosine": math.cos(w / (2 * cutoff)), "hamming": 0.54 + 0.46 * math.cos(w / cutoff),
              "hann": 0.5 + 0.5 * math.cos(w / cutoff)}[ftype]
    assert H[m] == pytest.approx(base[m] * factor, rel=1e-12)


def test_weight_is_cosine_of_the_ray_angle():
    # R/sqrt(R^2 + kk^2 + pp^2) on isocentre-scaled centres = SDD/|ray to the physical cell centre|.
    g = P.PixelGeometry(make_cone_params(DetectorWidth=7, DetectorHeight=6))
    W = P.weights(g)
    for k in range(6):
        for j in range(7):
            u = (j - 3.0) * g.du
            v = (k - 2.5) * g.dv
            assert W[k, j] == pytest.approx(g.sdd / math.sqrt(g.sdd ** 2 + u * u + v * v), rel=1e-13)


@pytest.mark.parametrize("nu", [9, 13, 16])
def test_filter_equals_direct_convolution_with_kak_slaney_taps(nu):
    # cutoff 1 masks nothing; the Ram-Lak |DFT| is then the DFT of the zero-phase taps
    # 2h = (1/(4 tau^2), -1/(pi^2 n^2 tau^2) at odd n), so filtering = linear convolution
    # of the flipped, weighted rows with those textbook taps.
    g = P.PixelGeometry(make_cone_params(DetectorWidth=nu, DetectorHeight=3, NumberOfViews=2))
    Hc, _ = closed_form_spectrum(g.zp, g.tau)
    assert np.all(1.0 / (8 * g.tau ** 2) + 2 * np.cos(np.outer(
        2 * np.pi * np.arange(g.zp) / g.zp, np.arange(1, g.zp // 2, 2))) @ (
        -1.0 / (2 * np.pi ** 2 * g.tau ** 2 * np.arange(1, g.zp // 2, 2) ** 2)) >= 0)
    rng = np.random.default_rng(3)
    proj = rng.standard_normal((2, 3, nu))
    out = P.filter_projections(g, proj, "ram-lak", 1.0)

    def tap(n):
        n = abs(n)
        if n == 0:
            return 1.0 / (4 * g.tau ** 2)
        return -1.0 / (np.pi ** 2 * n ** 2 * g.tau ** 2) if n % 2 else 0.0

    W = P.weights(g)
    for v in range(2):
        for r in range(3):
            x = W[r] * proj[v, r, ::-1]
            ref = [sum(x[m] * tap(n - m) for m in range(nu)) for n in range(nu)]
            np.testing.assert_allclose(out[v, r], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_interpolation_matches_scipy_regular_grid_with_zero_fill():
    # The reference's CPU ConeBeam does this step with RegularGridInterpolator((p, ki), Q,
    # bounds_error=False, fill_value=0) (python/ConeBeam.py:143-144).
    g = P.PixelGeometry(make_cone_params(DetectorWidth=11, DetectorHeight=8))
    rng = np.random.default_rng(5)
    q = rng.standard_normal((8, 11))
    U = rng.uniform(g.ki[0] - 1.0, g.ki[-1] + 1.0, 4000)
    V = rng.uniform(g.p[0] - 1.0, g.p[-1] + 1.0, 4000)
    # grid nodes, the lower edges exactly and the upper edges from just inside (on the upper
    # edge itself (U - ki0)/dki rounds to either side of n-1, in the reference as here)
    U[:10], V[:10] = g.ki[:-1], g.p[3]
    U[11:19], V[11:19] = g.ki[-1] - 1e-9, g.p
    U[19:27], V[19:27] = g.ki[0], g.p
    U[27:37], V[27:37] = g.ki[:-1], g.p[-1] - 1e-9
    ref = RegularGridInterpolator((g.p, g.ki), q, bounds_error=False, fill_value=0.0)(
        np.stack([V, U], axis=-1))
    np.testing.assert_allclose(P.interpolate(q, g.ki, g.p, U, V), ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("theta", [0.0, 0.7, 2.0])
def test_hot_cell_lights_the_ray_from_the_source(theta):
    # A single non-zero Q cell at (ki[j], p[k]) backprojects onto the line from the source
    # R(-sin, cos, 0) through the isocentre-plane point ki[j](cos, sin, 0) + p[k] z.
    g = P.PixelGeometry(make_cone_params(ReconX=32, ReconY=32, ReconZ=32, NumberOfViews=1,
                                         AngleCoverage=1.0))
    g.angles = np.array([theta])
    j, k = 33, 9
    q = np.zeros((1, g.nv, g.nu))
    q[0, k, j] = 1.0
    vol = P.backproject(g, q, scale=1.0)
    S = g.R * np.array([-math.sin(theta), math.cos(theta), 0.0])
    T = g.ki[j] * np.array([math.cos(theta), math.sin(theta), 0.0]) + np.array([0, 0, g.p[k]])
    d = (T - S) / np.linalg.norm(T - S)
    Z, Y, X = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    pts = np.stack([X, Y, Z], -1) - S
    dist = np.linalg.norm(pts - (pts @ d)[..., None] * d, axis=-1)
    lit = vol > 0.5 * vol.max()
    assert lit.sum() >= 8
    assert dist[lit].max() < 1.0 * (g.x[1] - g.x[0])


@pytest.mark.parametrize("centre", [(0.0, 0.0, 0.0), (2.0, -1.5, 1.0)])
def test_fdk_of_analytic_ball_recovers_density_and_position(centre):
    # Projections of a uniform ball by exact chords, in the reference's data convention
    # (column n of a view sits at ki[nu-1-n] once fliplr-ed).  ConeBeamGPU applies no tau and
    # no 1/2 over the full circle, so its image of density 1 is 2/tau.
    g = P.PixelGeometry(make_cone_params())
    ell = [(centre, (7.0, 7.0, 7.0), (0, 0, 0), 1.0)]
    proj = np.zeros((g.ns, g.nv, g.nu))
    kk = g.ki[::-1]
    for i, th in enumerate(g.angles):
        S = g.R * np.array([-math.sin(th), math.cos(th), 0.0])
        Pt = (kk[None, :, None] * np.array([math.cos(th), math.sin(th), 0.0])
              + g.p[:, None, None] * np.array([0.0, 0.0, 1.0]))
        proj[i] = analytic.line_integral(ell, S, Pt)
    rec = P.fdk(g, proj) * g.tau / 2.0
    Z, Y, X = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    r = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2)
    assert abs(rec[r < 4.0].mean() - 1.0) < 0.03
    assert abs(rec[r > 10.0].mean()) < 0.03
    m = np.clip(rec, 0, None)
    for axis, c in zip((X, Y, Z), centre):
        assert abs((m * axis).sum() / m.sum() - c) < 0.3


def test_edge_margin_flags_samples_on_the_bounds():
    g = P.PixelGeometry(make_cone_params(ReconX=16, ReconY=16, ReconZ=16, NumberOfViews=4))
    m = P.edge_margin(g)
    assert m.shape == (16, 16, 16)
    # the square grid's corners leave the inscribed field of view: some voxel crosses an edge
    assert np.isfinite(m).all() and m.min() < 5.0
