"""Torch-level operators over the HIP library: DD forward projection, FDK filter, backprojection.

Tensors stay on the GPU; the library enqueues its kernels on torch's current HIP stream and
returns.  Scratch memory (per-view tables, the z-prefix tables of the volume, twiddles) is a torch
uint8 tensor handed to the library as its workspace, so the caching allocator owns it.
Inputs must be float32, contiguous and on a ROCm device: there is no CPU path.
"""
from __future__ import annotations

import ctypes

import torch

from . import _native as N
from .geometry import ScanGeometry


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _require(t: torch.Tensor, shape, name: str, dtype=torch.float32) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor on a ROCm device")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must live on the GPU (got {t.device}); there is no CPU path")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


def _workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def dd_forward(vol: torch.Tensor, geom: ScanGeometry, out: torch.Tensor | None = None) -> torch.Tensor:
    """DD forward projection vol[nz,ny,nx] -> proj[nView,nv,nu] (overwritten).

    Replaces Reconstruction.forward/distance_forward (python/Reconstruction_pycuda.py:750-1059).
    """
    _require(vol, geom.vol_shape, "vol")
    dev = vol.device
    if out is None:
        out = torch.empty(geom.proj_shape, dtype=torch.float32, device=dev)
    _require(out, geom.proj_shape, "proj")
    lib = N.lib()
    g = geom.ctypes_struct()
    ang, br = geom.device_views(dev)
    nbytes = lib.ctr_dd_forward_workspace_bytes(ctypes.byref(g))
    ws = _workspace(nbytes, dev)
    N.check(lib.ctr_dd_forward(ctypes.byref(g), _ptr(ang), _ptr(br), _ptr(vol), _ptr(out),
                               _ptr(ws), nbytes, _stream(dev)), "ctr_dd_forward")
    return out


def filter_spectrum(geom: ScanGeometry, filter_type: str | None = None, cutoff: float | None = None,
                    device="cuda") -> torch.Tensor:
    """Filter magnitude in FFT order, float64[N] with N = zero-padded length, computed on the GPU.

    Replaces Reconstruction.Filter (python/Reconstruction_pycuda.py:186-221).
    """
    ftype = geom.filter_type if filter_type is None else filter_type
    cut = geom.cutoff if cutoff is None else float(cutoff)
    if ftype not in N.FILTER_CODES:
        raise ValueError(f"unknown FilterType {ftype!r}; expected one of {sorted(N.FILTER_CODES)}")
    lib = N.lib()
    g = geom.ctypes_struct()
    out = torch.empty(geom.zero_padded_length, dtype=torch.float64, device=device)
    N.check(lib.ctr_filter_spectrum(ctypes.byref(g), N.FILTER_CODES[ftype], cut, _ptr(out),
                                    _stream(out.device)), "ctr_filter_spectrum")
    return out


def fdk_filter(proj: torch.Tensor, geom: ScanGeometry, spectrum: torch.Tensor | None = None,
               out: torch.Tensor | None = None) -> torch.Tensor:
    """Cosine weight + ramp-family filter of every detector row (``Filtering()``,
    python/Reconstruction_pycuda.py:223-251).  ``out`` may be ``proj`` (in place)."""
    if proj.dim() != 3 or tuple(proj.shape[1:]) != (geom.nv, geom.nu):
        raise ValueError(f"proj has shape {tuple(proj.shape)}, expected [n, {geom.nv}, {geom.nu}]")
    _require(proj, None, "proj")
    dev = proj.device
    if spectrum is None:
        spectrum = filter_spectrum(geom, device=dev)
    _require(spectrum, (geom.zero_padded_length,), "spectrum", torch.float64)
    if out is None:
        out = torch.empty_like(proj)
    _require(out, tuple(proj.shape), "out")
    lib = N.lib()
    g = geom.ctypes_struct()
    nbytes = lib.ctr_fdk_filter_workspace_bytes(ctypes.byref(g))
    ws = _workspace(nbytes, dev)
    N.check(lib.ctr_fdk_filter(ctypes.byref(g), _ptr(spectrum), _ptr(proj), _ptr(out),
                               int(proj.shape[0]), _ptr(ws), nbytes, _stream(dev)), "ctr_fdk_filter")
    return out


def dd_backproject(proj: torch.Tensor, geom: ScanGeometry, mode: str = "ref", scale: float | None = None,
                   zero_tol: float = 0.0, out: torch.Tensor | None = None,
                   accumulate: bool = False) -> torch.Tensor:
    """DD backprojection proj[nView,nv,nu] -> vol[nz,ny,nx].

    ``mode='ref'``: the reference's FDK backprojector with its half-pixel detector grid
    (python/Reconstruction_pycuda.py:320-435, kernel python/GPUFuncs_pycuda.py:837-956);
    ``mode='edge'``: the same with the forward projector's cell edges; ``mode='adjoint'``: the
    exact transpose A^T of :func:`dd_forward` (no reference counterpart).  ``scale`` multiplies
    the result; it defaults to dtheta for the FDK modes (the reference's Q = proj * dtheta) and
    to 1 for the adjoint.
    """
    if mode not in N.BP_MODES:
        raise ValueError(f"mode must be one of {sorted(N.BP_MODES)}, got {mode!r}")
    _require(proj, geom.proj_shape, "proj")
    if mode == "adjoint":
        geom.check_supported()
    dev = proj.device
    if out is None:
        out = torch.empty(geom.vol_shape, dtype=torch.float32, device=dev)
        accumulate = False
    _require(out, geom.vol_shape, "vol")
    lib = N.lib()
    g = geom.ctypes_struct()
    ang, br = geom.device_views(dev)
    m = N.BP_MODES[mode]
    nbytes = lib.ctr_dd_backproject_workspace_bytes(ctypes.byref(g), m)
    ws = _workspace(nbytes, dev)
    if scale is None:
        scale = 1.0 if mode == "adjoint" else geom.dtheta
    N.check(lib.ctr_dd_backproject(ctypes.byref(g), _ptr(ang), _ptr(br), _ptr(proj), _ptr(out), m,
                                   float(scale), float(zero_tol), int(bool(accumulate)), _ptr(ws),
                                   nbytes, _stream(dev)), "ctr_dd_backproject")
    return out


def dd_adjoint(proj: torch.Tensor, geom: ScanGeometry, out: torch.Tensor | None = None) -> torch.Tensor:
    """A^T proj: exact transpose of :func:`dd_forward` (matched backprojector)."""
    return dd_backproject(proj, geom, mode="adjoint", out=out)


def fdk(proj: torch.Tensor, geom: ScanGeometry, mode: str = "ref",
        spectrum: torch.Tensor | None = None) -> torch.Tensor:
    """Filtering() then backward() (python/FDK_512.py:31-33), all on the GPU."""
    q = fdk_filter(proj, geom, spectrum)
    return dd_backproject(q, geom, mode=mode)
