"""ctypes front end of the C restatement of the pixel oracle's backprojection
(``c/pixel_backproject.c``, in the library :mod:`oracle.dd_c` builds) and a multi-threaded
restatement of its filter.

TEST INFRASTRUCTURE: used by the tests and the benchmark's CPU-baseline leg as the checker and
as the CPU time of a compiled port; the product path never imports it.

``backproject`` returns what :func:`oracle.pixel_oracle.backproject` returns and, with
``margin=True``, what :func:`oracle.pixel_oracle.edge_margin` returns over the same views (the C
loop performs the same FP64 operations in the same order, on the oracle's cell centres, voxel
coordinates and angles), split over OpenMP threads.  ``filter_projections`` is
:func:`oracle.pixel_oracle.filter_projections` with its FFTs batched over views on ``scipy.fft``
workers.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import dd_c
from . import pixel_oracle as PO


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def backproject(g: PO.PixelGeometry, q: np.ndarray, views, scale: float | None = None,
                threads: int | None = None, margin: bool = False):
    """Sum over ``views`` of ``scale * W * bilinear(q[n]; U, V)`` -> [nz, ny, nx] float64, where
    ``q[n]`` is the filtered projection of view ``views[n]`` in the oracle's mirrored column order;
    with ``margin`` also the per-voxel edge margin over those views."""
    views = [int(v) for v in views]
    threads = dd_c.host_threads() if threads is None else int(threads)
    scale = g.dtheta if scale is None else float(scale)
    qs = np.ascontiguousarray(q, dtype=np.float64)
    if qs.shape != (len(views), g.nv, g.nu):
        raise ValueError(f"projections {qs.shape} != ({len(views)} views, {g.nv}, {g.nu})")
    cos_a = np.asarray([math.cos(g.angles[i]) for i in views], dtype=np.float64)
    sin_a = np.asarray([math.sin(g.angles[i]) for i in views], dtype=np.float64)
    ki, p, x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (g.ki, g.p, g.x, g.y, g.z))
    vol = np.zeros((g.nz, g.ny, g.nx), dtype=np.float64)
    m = np.full((g.nz, g.ny, g.nx), np.inf) if margin else None
    rc = dd_c.lib().pxo_backproject(
        _dp(qs), len(views), g.nv, g.nu, _dp(cos_a), _dp(sin_a), g.R, scale, _dp(ki), _dp(p),
        _dp(x), _dp(y), _dp(z), g.nx, g.ny, g.nz, threads, _dp(vol),
        _dp(m) if margin else None)
    if rc != 0:
        raise RuntimeError(f"pxo_backproject failed with code {rc}")
    return (vol, m) if margin else vol


def filter_projections(g: PO.PixelGeometry, proj: np.ndarray, filter_type: str = "ram-lak",
                       cutoff: float = 0.3, threads: int | None = None, chunk: int = 8) -> np.ndarray:
    """Weight, mirror and row-filter each view (as :func:`oracle.pixel_oracle.filter_projections`),
    ``chunk`` views per batched FFT; float64 out."""
    import scipy.fft as sf

    threads = dd_c.host_threads() if threads is None else int(threads)
    Hs = np.fft.ifftshift(PO.spectrum(g.zp + 1, g.tau, filter_type, cutoff))
    W = PO.weights(g)
    proj = np.asarray(proj)
    out = np.empty(proj.shape, dtype=np.float64)
    for i0 in range(0, proj.shape[0], chunk):
        wp = W[None] * np.asarray(proj[i0:i0 + chunk, :, ::-1], dtype=np.float64)
        f = sf.fft(wp, g.zp, axis=-1, workers=threads)
        f *= Hs
        out[i0:i0 + chunk] = sf.ifft(f, axis=-1, workers=threads).real[..., :g.nu]
    return out
