"""Pins the oracle's FDK filter (python/Reconstruction_pycuda.py:186-251) without reusing its code path."""
import numpy as np
import pytest

from oracle import dd_oracle as O
from tests.conftest import make_params


def closed_form_spectrum(zp, tau):
    """|DFT| of the discrete Ram-Lak taps as a cosine series (no FFT).

    Taps live on x = -zp/2 .. zp/2-1: h(0) = 1/(8 tau^2), h(odd n) = -1/(2 pi^2 tau^2 n^2),
    h(even n != 0) = 0.  Because h is even, |DFT|(w) = |h0 + 2 sum_{n odd > 0} h(n) cos(n w)|
    with w = 2 pi m / zp; fftshift puts m = -zp/2 at index 0.
    """
    m = np.arange(zp) - zp // 2
    w = 2 * np.pi * m / zp
    n = np.arange(1, zp // 2, 2)
    hn = -1.0 / (2 * np.pi ** 2 * tau ** 2 * n ** 2)
    H = np.abs(1.0 / (8 * tau ** 2) + 2.0 * np.cos(np.outer(w, n)) @ hn)
    return H, w


@pytest.mark.parametrize("zp", [8, 64, 1024])
@pytest.mark.parametrize("tau", [0.2666, 0.6667, 1.3])
def test_ramlak_spectrum_matches_cosine_series(zp, tau):
    H = O.ramp_spectrum(zp + 1, tau, "ram-lak", 1.0)
    Hc, w = closed_form_spectrum(zp, tau)
    Hc[np.abs(w) > np.pi / (2 * tau)] = 0  # frequency mask in rad/sample vs mm (Q4)
    np.testing.assert_allclose(H, Hc, rtol=0, atol=1e-12 * Hc.max())


def test_ramlak_tends_to_ideal_ramp():
    # Continuum limit: the discrete Ram-Lak magnitude approaches |w|/(2 pi tau^2) * (1/2) ... the
    # reference's taps carry the 1/2 full-scan redundancy factor, so H(w) ~ |w| / (4 pi tau^2).
    zp, tau = 4096, 0.5
    H = O.ramp_spectrum(zp + 1, tau, "ram-lak", 1.0)
    w = 2 * np.pi * (np.arange(zp) - zp // 2) / zp
    sel = (np.abs(w) > 0.05) & (np.abs(w) < 1.0)
    np.testing.assert_allclose(H[sel], np.abs(w[sel]) / (4 * np.pi * tau ** 2), rtol=2e-3)


@pytest.mark.parametrize("ftype,at_nyquist", [
    ("ram-lak", 1.0), ("shepp-logan", 2 / np.pi), ("cosine", 0.0), ("hamming", 0.08), ("hann", 0.0)])
def test_window_value_at_nyquist(ftype, at_nyquist):
    # Textbook window values at w = pi with cutoff 1: sinc(1/2)=2/pi, cos(pi/2)=0, 0.54-0.46, 0.5-0.5.
    zp, tau = 64, 0.1  # small tau: the frequency mask is inactive
    H = O.ramp_spectrum(zp + 1, tau, ftype, 1.0)
    Hr = O.ramp_spectrum(zp + 1, tau, "ram-lak", 1.0)
    assert H[0] == pytest.approx(Hr[0] * at_nyquist, abs=1e-12 * Hr.max())


def test_shepp_logan_zeroes_dc():
    # Q3: the reference replaces the DC bin by H[DC]*sin(0) = 0.
    zp = 128
    H = O.ramp_spectrum(zp + 1, 0.3, "shepp-logan", 1.0)
    assert H[zp // 2] == 0.0
    Hr = O.ramp_spectrum(zp + 1, 0.3, "ram-lak", 1.0)
    assert Hr[zp // 2] > 0.0


def test_frequency_mask_active_for_coarse_tau():
    # Q4 at C2-like sampling (tau = 1.0*1000/1500): bins with |w| > pi/(2 tau) are zero.
    zp, tau = 1024, 1000.0 / 1500.0
    H = O.ramp_spectrum(zp + 1, tau, "shepp-logan", 1.0)
    w = 2 * np.pi * (np.arange(zp) - zp // 2) / zp
    assert np.all(H[np.abs(w) > np.pi / (2 * tau)] == 0)
    assert np.all(H[(np.abs(w) < np.pi / (2 * tau) - 0.01) & (w != 0)] > 0)


def test_zero_padded_length():
    assert O.zero_padded_length(384) == 1024
    assert O.zero_padded_length(512) == 1024
    assert O.zero_padded_length(513) == 1024
    assert O.zero_padded_length(1024) == 2048
    assert O.zero_padded_length(256) == 512


def test_cosine_weight_is_one_at_central_edge():
    # Q2: the pre-weight is evaluated at cell edges (j - nu/2) du, so for even nu, nv the
    # cell (nv/2, nu/2) sits at edge coordinate (0, 0) and its weight is exactly 1.
    g = O.Geometry(make_params(NumberOfDetectorPixels=[20, 14]))
    W = O.fdk_weights(g)
    assert W[7, 10] == 1.0
    assert np.all(W <= 1.0)


@pytest.mark.parametrize("ftype", O.FILTER_TYPES)
def test_filtering_equals_direct_linear_convolution(ftype):
    # Zero-padded FFT filtering with a real even spectrum == linear convolution with the
    # spatial kernel g = ifft(H); checked by an O(n^2) loop, not by FFTs of the data.
    g = O.Geometry(make_params(NumberOfDetectorPixels=[13, 3], NumberOfViews=2, FilterType=ftype,
                               cutoff=0.8))
    rng = np.random.default_rng(1)
    proj = rng.standard_normal((2, 3, 13)).astype(np.float32)
    out = O.filter_projections(g, proj)
    zp = O.zero_padded_length(13)
    tau = g.du * g.sad / g.sdd
    H = O.ramp_spectrum(zp + 1, tau, ftype, 0.8)
    kern = np.real(np.fft.ifft(np.fft.ifftshift(H)))
    W = O.fdk_weights(g)
    ref = np.zeros((2, 3, 13))
    for v in range(2):
        for r in range(3):
            x = W[r] * proj[v, r].astype(np.float64)
            for n in range(13):
                ref[v, r, n] = tau * sum(x[m] * kern[(n - m) % zp] for m in range(13))
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-6 * np.abs(ref).max())
