"""ctypes binding of the HIP library ``lib/libctrecon.so`` (C ABI in ``include/ctrecon.h``).

The library is built in-tree by :func:`build` (``hipcc --offload-arch=gfx950``).  Loading it
imports torch first so that its HIP runtime (``libamdhip64.so.7``) is the one the library
binds to: device pointers and streams handed over from torch tensors are then valid in the
library.  There is no CPU fallback: a missing library raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libctrecon.so")
SOURCES = [os.path.join(PKG_DIR, "csrc", "ctrecon.hip"), os.path.join(PKG_DIR, "csrc", "prior.hip"),
           os.path.join(PKG_DIR, "csrc", "pixel.hip")]
INCLUDE_DIR = os.path.join(REPO_DIR, "include")
# instruction-issue microbenchmarks for bench.py --calibrate-valu (not part of the product ABI)
CALIB_PATH = os.path.join(LIB_DIR, "libctrcalib.so")
CALIB_SOURCES = [os.path.join(PKG_DIR, "csrc", "calibrate.hip")]

# Exported symbols (every function declared in include/ctrecon.h).
EXPORTS = (
    "ctr_version", "ctr_strerror", "ctr_select_branches", "ctr_zero_padded_length",
    "ctr_dd_forward_workspace_bytes", "ctr_dd_forward",
    "ctr_fdk_filter_workspace_bytes", "ctr_filter_spectrum", "ctr_fdk_filter",
    "ctr_dd_backproject_workspace_bytes", "ctr_dd_backproject", "ctr_dd_backproject_static",
    "ctr_pixel_filter_spectrum", "ctr_pixel_filter",
    "ctr_pixel_backproject_workspace_bytes", "ctr_pixel_backproject",
    "ctr_conv3x3_bf16_forward", "ctr_conv3x3_bf16_wgrad_workspace_bytes", "ctr_conv3x3_bf16_wgrad",
)

CTR_OK = 0
CTR_ERR_INVALID = 1
CTR_ERR_UNSUPPORTED = 2
CTR_ERR_WORKSPACE = 3
CTR_ERR_HIP = 4
CTR_ERR_SCRATCH = 5

CTR_DETECTOR_FLAT = 0
CTR_DETECTOR_CURVED = 1
CTR_BRANCH_X, CTR_BRANCH_Y, CTR_BRANCH_Z = 0, 1, 2
FILTER_CODES = {"ram-lak": 0, "shepp-logan": 1, "cosine": 2, "hamming": 3, "hann": 4}
BP_MODES = {"ref": 0, "edge": 1, "adjoint": 2}


class CtrGeometry(ctypes.Structure):
    """Mirror of ``struct ctr_geometry`` (include/ctrecon.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("nu", ctypes.c_int32), ("nv", ctypes.c_int32), ("n_views", ctypes.c_int32),
        ("detector_shape", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("du", ctypes.c_double), ("dv", ctypes.c_double),
        ("sad", ctypes.c_double), ("sdd", ctypes.c_double),
        ("source_init", ctypes.c_double * 3), ("detector_init", ctypes.c_double * 3),
        ("helical", ctypes.c_double), ("recon_center", ctypes.c_double * 3),
        ("det_offset", ctypes.c_double * 2),
    ]


class CtrError(RuntimeError):
    """Raised for a non-zero ``ctr_status``; replaces the reference's print + sys.exit()
    (python/Reconstruction_pycuda.py:378-385, 393-397)."""

    def __init__(self, code: int, where: str):
        self.code = code
        msg = _lib.ctr_strerror(code).decode() if _lib is not None else f"code {code}"
        super().__init__(f"{where}: {msg} (ctr_status {code})")


_lib = None


def hipcc_command(out: str = LIB_PATH, sources=None) -> list[str]:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return [hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            # device code never produces or tests NaNs: lets min/max skip NaN quieting
            "-Xarch_device", "-fno-honor-nans",
            "-I", INCLUDE_DIR, "-o", out, *(SOURCES if sources is None else sources)]


def _compile(out: str, sources, verbose: bool) -> str:
    tmp = out + ".tmp"
    cmd = hipcc_command(tmp, sources)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)
    if verbose:
        print(" ".join(cmd))
    os.replace(tmp, out)
    return out


def build(verbose: bool = False) -> str:
    """Compile the HIP library for gfx950 into ``lib/libctrecon.so`` (always recompiles), and the
    calibration microbenchmarks into ``lib/libctrcalib.so``."""
    os.makedirs(LIB_DIR, exist_ok=True)
    _compile(CALIB_PATH, CALIB_SOURCES, verbose)
    return _compile(LIB_PATH, SOURCES, verbose)


def _declare(lib):
    c_int, c_size, c_vp, c_dbl, c_flt = ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_double, ctypes.c_float
    gp = ctypes.POINTER(CtrGeometry)
    sig = {
        "ctr_version": (c_int, []),
        "ctr_strerror": (ctypes.c_char_p, [c_int]),
        "ctr_select_branches": (c_int, [gp, c_vp, c_vp]),
        "ctr_zero_padded_length": (c_int, [c_int]),
        "ctr_dd_forward_workspace_bytes": (c_size, [gp]),
        "ctr_dd_forward": (c_int, [gp, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
        "ctr_fdk_filter_workspace_bytes": (c_size, [gp]),
        "ctr_filter_spectrum": (c_int, [gp, c_int, c_dbl, c_vp, c_vp]),
        "ctr_fdk_filter": (c_int, [gp, c_vp, c_vp, c_vp, c_int, c_vp, c_size, c_vp]),
        "ctr_dd_backproject_workspace_bytes": (c_size, [gp, c_int]),
        "ctr_dd_backproject": (c_int, [gp, c_vp, c_vp, c_vp, c_vp, c_int, c_flt, c_flt, c_int,
                                       c_vp, c_size, c_vp]),
        "ctr_dd_backproject_static": (c_int, [gp, c_vp, c_vp, c_vp, c_int, c_flt, c_flt, c_int, c_vp]),
        "ctr_pixel_filter_spectrum": (c_int, [c_int, c_dbl, c_int, c_dbl, c_vp, c_vp]),
        "ctr_pixel_filter": (c_int, [gp, c_vp, c_vp, c_vp, c_int, c_vp, c_size, c_vp]),
        "ctr_pixel_backproject_workspace_bytes": (c_size, [gp]),
        "ctr_pixel_backproject": (c_int, [gp, c_vp, c_vp, c_vp, c_flt, c_int, c_vp, c_size, c_vp]),
        "ctr_conv3x3_bf16_forward": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int,
                                             c_int, c_int, c_vp]),
        "ctr_conv3x3_bf16_wgrad_workspace_bytes": (c_size, [c_int, c_int, c_int, c_int, c_int]),
        "ctr_conv3x3_bf16_wgrad": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int,
                                           c_vp, c_size, c_vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Load (once) and return the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP library {LIB_PATH} is missing; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
        import torch  # noqa: F401  -- bind to torch's HIP runtime before loading ours
        handle = ctypes.CDLL(LIB_PATH)
        _declare(handle)
        _lib = handle
    return _lib


def check(code: int, where: str) -> None:
    if code != CTR_OK:
        raise CtrError(code, where)
