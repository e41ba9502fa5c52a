"""ctypes front end of the C restatements of the oracle's FDK backprojection (``c/dd_backproject.c``)
and DD forward projection (``c/dd_forward.c``), and a multi-threaded restatement of its filter;
the library also holds the pixel oracle's backprojection (``c/pixel_backproject.c``, front end in
:mod:`oracle.pixel_c`).

TEST INFRASTRUCTURE: used by the tests, the smoke check and the benchmark's CPU-baseline leg as
the checker and as the CPU time of a compiled port; the product path never imports it.

``backproject`` returns what :func:`oracle.dd_oracle.backproject` returns (the C loop performs the
same FP64 operations in the same order; it takes the detector grid, voxel centres and per-view
source heights from :mod:`oracle.dd_oracle`), split over OpenMP threads.  ``filter_projections``
is :func:`oracle.dd_oracle.filter_projections` with its FFTs batched over views and run by
``scipy.fft`` on several workers.  ``forward_project_view`` returns what
:func:`oracle.dd_oracle.forward_project_view` returns (the slab loop in C, the view's line tables
from :func:`oracle.dd_oracle.fp_view_tables`), its cells split over OpenMP threads.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import dd_oracle as O

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(ORACLE_DIR, "c", n) for n in ("dd_backproject.c", "dd_forward.c", "pixel_backproject.c")]
LIB_PATH = os.path.join(ORACLE_DIR, "lib", "liboracle.so")
CFLAGS = ["-O3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wextra"]

_lib = None


def build() -> str:
    """Compile ``c/*.c`` into ``lib/liboracle.so`` with gcc (always recompiles)."""
    global _lib
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    subprocess.run(["gcc", *CFLAGS, *SOURCES, "-o", tmp, "-lm"], check=True)
    os.replace(tmp, LIB_PATH)
    _lib = None
    return LIB_PATH


def lib():
    """The loaded C oracle; raises ``RuntimeError`` when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run oracle.dd_c.build()")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.ddo_backproject.restype = ctypes.c_int
        L.ddo_backproject.argtypes = [
            ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int, ctypes.c_int,
            dp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
            dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp]
        i64 = ctypes.c_int64
        L.ddo_forward_view.restype = ctypes.c_int
        L.ddo_forward_view.argtypes = [
            ctypes.POINTER(ctypes.c_float), i64, i64, i64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            dp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            dp, dp, dp, dp, dp, dp, dp, dp, dp, i64, ctypes.c_double, ctypes.c_int, dp]
        L.pxo_backproject.restype = ctypes.c_int
        L.pxo_backproject.argtypes = [
            dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_double,
            dp, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        _lib = L
    return _lib


def host_threads() -> int:
    return len(os.sched_getaffinity(0))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def backproject(g: O.Geometry, q: np.ndarray, mode: str = "ref", zero_tol: float | None = None,
                views=None, threads: int | None = None, out: np.ndarray | None = None) -> np.ndarray:
    """Sum over ``views`` (default: every view of the scan) of the FDK-weighted DD backprojection
    of ``q[n] * dtheta`` -> [nz, ny, nx] float64; ``q[n]`` is the filtered projection of view
    ``views[n]``.  With ``out`` the views are added into it."""
    views = list(range(g.nview)) if views is None else [int(v) for v in views]
    threads = host_threads() if threads is None else int(threads)
    ki, p, du, dv = O.bp_detector_grid(g, mode)
    ki = np.ascontiguousarray(ki, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    xc, yc, zc = (np.ascontiguousarray(a, dtype=np.float64) for a in g.centres())
    th = np.asarray([g.angles[i] for i in views], dtype=np.float64)
    cos_th = np.asarray([np.cos(t) for t in th], dtype=np.float64)
    sin_th = np.asarray([np.sin(t) for t in th], dtype=np.float64)
    src_z = np.asarray([g.source_z0 + g.helical * t / (2 * O.PI) for t in th], dtype=np.float64)
    qs = np.ascontiguousarray(q, dtype=np.float32)
    if qs.shape != (len(views), g.nv, g.nu):
        raise ValueError(f"projections {qs.shape} != ({len(views)} views, {g.nv}, {g.nu})")
    if out is None:
        out = np.zeros((g.nz, g.ny, g.nx), dtype=np.float64)
    elif out.shape != (g.nz, g.ny, g.nx) or out.dtype != np.float64 or not out.flags.c_contiguous:
        raise ValueError("out must be a contiguous float64 [nz, ny, nx] array")
    rc = lib().ddo_backproject(
        qs.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(views), g.nv, g.nu,
        _dp(cos_th), _dp(sin_th), _dp(src_z), float(g.source[0]), float(g.source[1]),
        float(g.detector[1]), float(g.sad), float(g.dtheta), _dp(ki), _dp(p), float(du), float(dv),
        int(g.shape == "Curved"), _dp(xc), _dp(yc), _dp(zc), g.nx, g.ny, g.nz,
        g.dx / 2.0, -g.dy / 2.0, -g.dz / 2.0, 0.0 if zero_tol is None else float(zero_tol), threads,
        _dp(out))
    if rc != 0:
        raise RuntimeError(f"ddo_backproject failed with code {rc}")
    return out


def forward_project_view(g: O.Geometry, vol: np.ndarray, i: int, zero_tol: float | None = 1e-6,
                         threads: int | None = None) -> np.ndarray:
    """View ``i`` of the DD forward projection of ``vol`` [nz, ny, nx] -> [nv, nu] float64, as
    :func:`oracle.dd_oracle.forward_project_view` (FP64; the volume's float32 values)."""
    threads = host_threads() if threads is None else int(threads)
    v = np.ascontiguousarray(vol, dtype=np.float32)
    if v.shape != (g.nz, g.ny, g.nx):
        raise ValueError(f"volume {v.shape} != ({g.nz}, {g.ny}, {g.nx})")
    br, *tabs = O.fp_view_tables(g, i)
    tabs = [np.ascontiguousarray(np.broadcast_to(t, (g.nv, g.nu)), dtype=np.float64).ravel() for t in tabs]
    sa, ta, va = O.AXES[br]
    strides = (1, g.nx, g.nx * g.ny)
    planes = [np.ascontiguousarray(p, dtype=np.float64) for p in g.planes()]
    dvec = (g.dx, -g.dy, -g.dz)
    ps, pt, pv = planes[sa], planes[ta], planes[va]
    out = np.zeros(g.nv * g.nu, dtype=np.float64)
    rc = lib().ddo_forward_view(
        v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), strides[sa], strides[ta], strides[va],
        len(ps) - 1, len(pt) - 1, len(pv) - 1, _dp(ps), _dp(pt), _dp(pv),
        float(dvec[sa]), float(dvec[ta]), float(dvec[va]), *[_dp(t) for t in tabs], g.nv * g.nu,
        0.0 if zero_tol is None else float(zero_tol), threads, _dp(out))
    if rc != 0:
        raise RuntimeError(f"ddo_forward_view failed with code {rc}")
    return out.reshape(g.nv, g.nu)


def forward_project(g: O.Geometry, vol: np.ndarray, views=None, zero_tol: float | None = 1e-6,
                    threads: int | None = None) -> np.ndarray:
    """``views`` (default all) of the DD forward projection -> [len(views), nv, nu] float64."""
    views = list(range(g.nview)) if views is None else [int(x) for x in views]
    return np.stack([forward_project_view(g, vol, i, zero_tol, threads) for i in views])


def filter_projections(g: O.Geometry, proj: np.ndarray, threads: int | None = None,
                       chunk: int = 8) -> np.ndarray:
    """Cosine weight, zero-padded FFT convolution with the window's ramp, crop, * tau
    (as :func:`oracle.dd_oracle.filter_projections`), ``chunk`` views per batched FFT."""
    import scipy.fft as sf

    threads = host_threads() if threads is None else int(threads)
    nu = g.nu
    zp = O.zero_padded_length(nu)
    R = g.sad
    D = g.sdd - R
    tau = g.du * R / (R + D)
    Hs = np.fft.ifftshift(O.ramp_spectrum(zp + 1, tau, g.filter_type, g.cutoff))
    W = O.fdk_weights(g)
    proj = np.asarray(proj)
    out = np.empty(proj.shape, dtype=np.float32)
    for i0 in range(0, proj.shape[0], chunk):
        wp = W[None] * np.asarray(proj[i0:i0 + chunk], dtype=np.float64)
        f = sf.fft(wp, zp, axis=-1, workers=threads)
        f *= Hs
        qf = sf.ifft(f, axis=-1, workers=threads).real
        out[i0:i0 + chunk] = (qf[..., :nu] * tau).astype(np.float32)
    return out
