"""FP64 NumPy oracle for the pixel-driven (voxel-driven, bilinear) cone-beam FDK.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``ctreconstruction_amd``) never calls it.

It restates ``ConeBeam.Reconstruction`` of python/ConeBeamGPU.py:203-321 (the
reference's pixel-driven GPU FDK), in double precision:

* geometry: ``fov``/``fovz`` from the detector size, voxel centres
  ``linspace(-fov/2, fov/2, n)``, angles ``linspace(0, AngleCoverage, ns+1)[:-1]``
  and ``dtheta`` (python/ConeBeamGPU.py:204-227);
* detector grid at the isocentre: ``ki = (j - (nx-1)/2) du R/(R+D)``,
  ``p = (k - (ny-1)/2) dv R/(R+D)`` (python/ConeBeamGPU.py:229-242);
* filter: ``ConeBeam.Filter`` (python/ConeBeamGPU.py:355-384) -- the Ram-Lak taps of
  ``Reconstruction.Filter`` times 2, the frequency mask ``|w| > pi*cutoff``; the scan
  uses ``ram-lak`` with ``cutoff = 0.3`` (python/ConeBeamGPU.py:236-239);
* pre-weight ``R/sqrt(R^2 + kk^2 + pp^2)`` at cell centres, ``fliplr`` of each view,
  zero-padded FFT filtering cropped to ``nx``, no ``tau`` factor
  (python/ConeBeamGPU.py:243, 296-302);
* backprojection: ``CalculateInterpolGrid`` (python/ConeBeamGPU.py:116-129) gives
  ``U = R t/(R-s)``, ``V = R z/(R-s)``, ``W = R^2/(R-s)^2`` with
  ``t = x cos + y sin``, ``s = -x sin + y cos`` (python/ConeBeamGPU.py:303-304);
  ``Interpol2dLineargpu`` (python/ConeBeamGPU.py:65-111) interpolates bilinearly on
  ``(ki, p)``, adds ``S * dtheta * W`` and adds ``fill_value = 0`` outside the grid.

Volume layout [nz][ny][nx] (the reference's flat ``tid = tz*X*Y + ty*X + tx``),
projections [n_views][nv][nu] (one ``[DetectorHeight, DetectorWidth]`` image per view).

Pinning lives in ``tests/test_oracle_pixel.py``: the spectrum against the cosine
series of the taps, the filter against a direct linear convolution with the textbook
Kak-Slaney taps, the interpolation against scipy's ``RegularGridInterpolator`` (the
routine the reference's CPU ``ConeBeam`` uses for the same step, python/ConeBeam.py:143-144),
a hot detector cell against the ray through it, and the FDK of analytic ball
projections against the ball's density and position.
"""
from __future__ import annotations

import math

import numpy as np

PI = np.pi
FILTER_TYPES = ("ram-lak", "shepp-logan", "cosine", "hamming", "hann")


class PixelGeometry:
    """Scalars of ``ConeBeam.Reconstruction`` (python/ConeBeamGPU.py:204-242).

    ``params`` is the ``ConeBeam`` params dict (AngleCoverage in radians), as
    python/ConeBeamGPU.py:164-197 reads it from the geometry text file."""

    def __init__(self, params: dict):
        self.R = float(params["SourceToAxis"])
        self.sdd = float(params["SourceToDetector"])
        self.D = self.sdd - self.R
        self.nu = int(params["DetectorWidth"])
        self.nv = int(params["DetectorHeight"])
        self.ns = int(params["NumberOfViews"])
        self.nx, self.ny, self.nz = int(params["ReconX"]), int(params["ReconY"]), int(params["ReconZ"])
        self.du = float(params["DetectorPixelWidth"])
        self.dv = float(params["DetectorPixelHeight"])
        self.coverage = float(params["AngleCoverage"])
        R, D = self.R, self.D
        width, height = self.nu * self.du, self.nv * self.dv
        self.fov = 2 * R * math.sin(math.atan(width / 2 / (D + R)))
        self.fovz = 2 * R * math.sin(math.atan(height / 2 / (D + R)))
        self.x = np.linspace(-self.fov / 2, self.fov / 2, self.nx)
        self.y = np.linspace(-self.fov / 2, self.fov / 2, self.ny)
        self.z = np.linspace(-self.fovz / 2, self.fovz / 2, self.nz)
        ang = np.linspace(0, self.coverage, self.ns + 1)
        self.angles = ang[0:-1]
        self.dtheta = ang[1] - ang[0]
        self.zp = int(2 ** (math.ceil(math.log2(2 * (self.nu - 1)))))
        self.ki = (np.arange(0, self.nu) - (self.nu - 1) / 2) * self.du * R / (R + D)
        self.p = (np.arange(0, self.nv) - (self.nv - 1) / 2) * self.dv * R / (R + D)
        self.tau = self.du * R / (D + R)


def spectrum(N: int, tau: float, filter_type: str = "ram-lak", cutoff: float = 0.3) -> np.ndarray:
    """``ConeBeam.Filter(N, pixel_size, FilterType, cutoff)`` (python/ConeBeamGPU.py:355-384).

    Centred order, length N-1.  Differs from ``Reconstruction.Filter`` by the factor 2 on
    ``|fftshift(fft(h))|`` and by the mask ``|w| > pi*cutoff`` (w in rad/sample)."""
    if filter_type not in FILTER_TYPES:
        raise ValueError(f"unknown FilterType {filter_type!r}")
    x = np.arange(0, N) - (N - 1) / 2
    h = np.zeros(len(x))
    h[np.where(x == 0)] = 1 / (8 * tau ** 2)
    odds = np.where(x % 2 == 1)
    h[odds] = -0.5 / (PI * tau * x[odds]) ** 2
    h = h[0:-1]
    H = np.abs(np.fft.fftshift(np.fft.fft(h))) * 2
    w = 2 * PI * x[0:-1] / (N - 1)
    if filter_type == "shepp-logan":
        zero = w == 0
        arg = w / (2 * cutoff)
        with np.errstate(divide="ignore", invalid="ignore"):
            H = H * np.where(zero, 0.0, np.sin(arg) / np.where(zero, 1.0, arg))
    elif filter_type == "cosine":
        H = H * np.cos(w / (2 * cutoff))
    elif filter_type == "hamming":
        H = H * (0.54 + 0.46 * np.cos(w / cutoff))
    elif filter_type == "hann":
        H = H * (0.5 + 0.5 * np.cos(w / cutoff))
    H[np.where(np.abs(w) > PI * cutoff)] = 0
    return H


def weights(g: PixelGeometry) -> np.ndarray:
    """``R / sqrt(R^2 + kk^2 + pp^2)`` on the centre grid, [nv, nu] (python/ConeBeamGPU.py:240-243)."""
    kk, pp = np.meshgrid(g.ki, g.p)
    return g.R / np.sqrt(g.R ** 2 + kk ** 2 + pp ** 2)


def filter_projections(g: PixelGeometry, proj: np.ndarray, filter_type: str = "ram-lak",
                       cutoff: float = 0.3) -> np.ndarray:
    """Per view: ``weight * fliplr(proj)``, row FFT filtering, crop (python/ConeBeamGPU.py:296-302).

    ``proj`` is [n_views][nv][nu]; the result Q is float64 in the reference's mirrored column
    order (column j of Q sits at ``ki[j]``)."""
    H = spectrum(g.zp + 1, g.tau, filter_type, cutoff)
    Hs = np.fft.ifftshift(H)
    W = weights(g)
    out = np.empty(proj.shape, dtype=np.float64)
    for i in range(proj.shape[0]):
        wp = W * np.fliplr(np.asarray(proj[i], dtype=np.float64))
        q = np.real(np.fft.ifft(Hs * np.fft.fft(wp, g.zp, axis=-1), axis=-1))
        out[i] = q[:, : g.nu]
    return out


def interpolate(q: np.ndarray, ki: np.ndarray, p: np.ndarray, U: np.ndarray, V: np.ndarray) -> np.ndarray:
    """``Interpol2dLineargpu``'s sample (python/ConeBeamGPU.py:80-102), vectorised, FP64.

    ``XLow = floor``, ``XHigh = ceil`` of the fractional index; the weights are the distances
    to the two grid nodes, with ``w1 = 1`` when both vanish (a sample on a node); a sample
    with ``XLow < 0`` or ``XHigh >= n`` (likewise in y) is 0 (``fill_value``)."""
    nu, nv = len(ki), len(p)
    fu = (U - ki[0]) / (ki[1] - ki[0])
    fv = (V - p[0]) / (p[1] - p[0])
    xl, xh = np.floor(fu), np.ceil(fu)
    yl, yh = np.floor(fv), np.ceil(fv)
    ok = (xl >= 0) & (xh < nu) & (yl >= 0) & (yh < nv)
    xl = np.where(ok, xl, 0).astype(np.int64)
    xh = np.where(ok, xh, 0).astype(np.int64)
    yl = np.where(ok, yl, 0).astype(np.int64)
    yh = np.where(ok, yh, 0).astype(np.int64)
    w1 = U - ki[xl]
    w2 = ki[xh] - U
    w3 = V - p[yl]
    w4 = p[yh] - V
    w1 = np.where((np.abs(w1) < 1e-4) & (np.abs(w2) < 1e-4), 1.0, w1)
    w3 = np.where((np.abs(w3) < 1e-4) & (np.abs(w4) < 1e-4), 1.0, w3)
    with np.errstate(divide="ignore", invalid="ignore"):  # out-of-grid samples are discarded
        M = (w2 / (w1 + w2)) * q[yl, xl] + (w1 / (w1 + w2)) * q[yl, xh]
        N = (w2 / (w1 + w2)) * q[yh, xl] + (w1 / (w1 + w2)) * q[yh, xh]
        S = (w4 / (w3 + w4)) * M + (w3 / (w3 + w4)) * N
    return np.where(ok, S, 0.0)


def sample_grid(g: PixelGeometry, angle: float):
    """``CalculateInterpolGrid`` for one view (python/ConeBeamGPU.py:116-129, 303-309).

    Returns U, V, W of shape [nz][ny][nx]."""
    xx, yy = np.meshgrid(g.x, g.y)
    t = xx * math.cos(angle) + yy * math.sin(angle)
    s = -xx * math.sin(angle) + yy * math.cos(angle)
    R = g.R
    U = np.broadcast_to(R * t / (R - s), (g.nz, g.ny, g.nx))
    V = R * g.z[:, None, None] / (R - s)[None]
    W = np.broadcast_to(R * R / ((R - s) * (R - s)), (g.nz, g.ny, g.nx))
    return U, V, W


def backproject(g: PixelGeometry, q: np.ndarray, scale: float | None = None,
                views=None) -> np.ndarray:
    """Sum over views of ``scale * W * bilinear(q_view; U, V)`` (python/ConeBeamGPU.py:303-315).

    ``scale`` defaults to the reference's ``dtheta``; ``views`` restricts the sum (a bounded
    sample for timing)."""
    scale = g.dtheta if scale is None else scale
    vol = np.zeros((g.nz, g.ny, g.nx))
    for i in (range(g.ns) if views is None else views):
        U, V, W = sample_grid(g, g.angles[i])
        vol += interpolate(q[i], g.ki, g.p, U, V) * scale * W
    return vol


def edge_margin(g: PixelGeometry) -> np.ndarray:
    """Per voxel, the smallest distance (in detector cells, over all views) of its sample from
    the lines where the reference's bounds test flips (index 0 or n-1 in u or v).

    A float32 evaluation of U or V can land on the other side of such a line than this FP64
    one; the contribution then jumps by a whole edge sample.  Tests compare only voxels whose
    margin exceeds the float32 resolution of the index."""
    m = np.full((g.nz, g.ny, g.nx), np.inf)
    for i in range(g.ns):
        U, V, _ = sample_grid(g, g.angles[i])
        fu = (U - g.ki[0]) / (g.ki[1] - g.ki[0])
        fv = (V - g.p[0]) / (g.p[1] - g.p[0])
        du = np.minimum(np.abs(fu), np.abs(fu - (g.nu - 1)))
        dv = np.minimum(np.abs(fv), np.abs(fv - (g.nv - 1)))
        m = np.minimum(m, np.minimum(du, dv))
    return m


def fdk(g: PixelGeometry, proj: np.ndarray, filter_type: str = "ram-lak", cutoff: float = 0.3) -> np.ndarray:
    """Filter + backprojection of ``ConeBeam.Reconstruction`` (python/ConeBeamGPU.py:292-315)."""
    return backproject(g, filter_projections(g, proj, filter_type, cutoff))
