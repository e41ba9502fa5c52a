"""Pixel-driven (voxel-driven, bilinear) cone-beam FDK: the reference's ``ConeBeam`` class of
python/ConeBeamGPU.py on the HIP library.

The reference reads the ``key: value`` geometry file, loads one ``.dat`` image per view and
reconstructs with per-view host filtering plus two CUDA kernels per view
(python/ConeBeamGPU.py:164-340).  Here the whole scan stays on the GPU: one filter launch for
all views (``ctr_pixel_filter``) and one backprojection launch (``ctr_pixel_backproject``)
that loops over the views inside the kernel.  There is no CPU path.

Conventions (python/ConeBeamGPU.py:204-243): voxel centres ``linspace(-fov/2, fov/2, n)`` with
``fov = 2 R sin(atan(W/2/SDD))`` (``fovz`` from the detector height), angles
``linspace(0, AngleCoverage, ns+1)[:-1]``, ``dtheta`` their step, detector cells at
``(j - (nu-1)/2) du`` scaled by ``R/SDD``, Ram-Lak with ``cutoff = 0.3``.  Volumes are
[nz][ny][nx] (the reference's flat ``tid = tz*X*Y + ty*X + tx``), projections
[n_views][nv][nu].
"""
from __future__ import annotations

import ctypes
import glob
import math
import warnings

import numpy as np
import torch

from . import _native as _nat
from .paramsfile import CONE_KEYS, ParamsFileError, field_of_view, read_params_file
from .projector import _ptr, _require, _stream, _workspace

# ConeBeamGPU's params dict adds ``Mode`` to ConeBeam's keys (python/ConeBeamGPU.py:164-168).
PIXEL_KEYS = CONE_KEYS + ("Mode",)
# LoadData's precisions (python/ConeBeamGPU.py:330-339).
PIXEL_PRECISIONS = {"float32": np.float32, "float64": np.float64, "int32": np.int32, "int64": np.int64}
DEFAULT_FILTER = "ram-lak"
DEFAULT_CUTOFF = 0.3


class PixelScan:
    """Geometry of one pixel-driven scan, from a ``ConeBeam`` params dict (AngleCoverage in radians)."""

    def __init__(self, params: dict):
        self.R = float(params["SourceToAxis"])
        self.sdd = float(params["SourceToDetector"])
        self.nu, self.nv = int(params["DetectorWidth"]), int(params["DetectorHeight"])
        self.du, self.dv = float(params["DetectorPixelWidth"]), float(params["DetectorPixelHeight"])
        self.n_views = int(params["NumberOfViews"])
        self.nx, self.ny, self.nz = int(params["ReconX"]), int(params["ReconY"]), int(params["ReconZ"])
        self.coverage = float(params["AngleCoverage"])
        if min(self.nu, self.nv) < 2:
            raise ValueError("the detector needs at least 2 x 2 cells (the reference indexes ki[1], p[1])")
        if self.n_views < 2:
            raise ValueError("NumberOfViews must be >= 2 (dtheta = angle[1] - angle[0])")
        if min(self.nx, self.ny, self.nz) < 1:
            raise ValueError("ReconVolume sizes must be positive")
        self.fov, self.fovz = field_of_view(params)
        ang = np.linspace(0, self.coverage, self.n_views + 1)
        self.angles = ang[:-1]
        self.dtheta = float(ang[1] - ang[0])
        self.zero_padded_length = int(2 ** math.ceil(math.log2(2 * (self.nu - 1))))
        self.tau = self.du * self.R / self.sdd

    @property
    def vol_shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def proj_shape(self):
        return (self.n_views, self.nv, self.nu)

    @staticmethod
    def _axis(n: int, extent: float) -> tuple[float, float]:
        """(spacing, centre offset) placing n voxel centres at linspace(-extent/2, extent/2, n);
        numpy puts a single sample at the start, -extent/2."""
        if n == 1:
            return 1.0, -extent / 2.0
        return extent / (n - 1), 0.0

    def ctypes_struct(self) -> _nat.CtrGeometry:
        g = _nat.CtrGeometry()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.nu, g.nv, g.n_views = self.nu, self.nv, self.n_views
        g.detector_shape = _nat.CTR_DETECTOR_FLAT
        (g.dx, cx), (g.dy, cy), (g.dz, cz) = (self._axis(self.nx, self.fov), self._axis(self.ny, self.fov),
                                              self._axis(self.nz, self.fovz))
        g.du, g.dv = self.du, self.dv
        g.sad, g.sdd = self.R, self.sdd
        g.recon_center[0], g.recon_center[1], g.recon_center[2] = cx, cy, cz
        return g

    def device_angles(self, device) -> torch.Tensor:
        return torch.as_tensor(self.angles, dtype=torch.float64, device=device)


def pixel_spectrum(n_fft: int, tau: float, filter_type: str = DEFAULT_FILTER,
                   cutoff: float = DEFAULT_CUTOFF, device="cuda") -> torch.Tensor:
    """``ConeBeam.Filter`` (python/ConeBeamGPU.py:355-384) in FFT order, float64[n_fft], on the GPU."""
    if filter_type not in _nat.FILTER_CODES:
        raise ValueError(f"unknown FilterType {filter_type!r}; expected one of {sorted(_nat.FILTER_CODES)}")
    if cutoff > 0.5 or cutoff < 0:
        # the reference prints ErrorDescription(4) and carries on (python/ConeBeamGPU.py:357-361)
        warnings.warn("Cutoff have to be pose between 0 and 0.5", stacklevel=2)
    out = torch.empty(int(n_fft), dtype=torch.float64, device=device)
    _nat.check(_nat.lib().ctr_pixel_filter_spectrum(int(n_fft), float(tau), _nat.FILTER_CODES[filter_type],
                                              float(cutoff), _ptr(out), _stream(out.device)),
            "ctr_pixel_filter_spectrum")
    return out


def pixel_filter(proj: torch.Tensor, scan: PixelScan, spectrum: torch.Tensor | None = None,
                 filter_type: str = DEFAULT_FILTER, cutoff: float = DEFAULT_CUTOFF,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """Weight, ``fliplr`` and row-filter every view (python/ConeBeamGPU.py:296-302).

    Returns Q [n, nv, nu] in the reference's mirrored column order."""
    if proj.dim() != 3 or tuple(proj.shape[1:]) != (scan.nv, scan.nu):
        raise ValueError(f"proj has shape {tuple(proj.shape)}, expected [n, {scan.nv}, {scan.nu}]")
    _require(proj, None, "proj")
    dev = proj.device
    if spectrum is None:
        spectrum = pixel_spectrum(scan.zero_padded_length, scan.tau, filter_type, cutoff, dev)
    _require(spectrum, (scan.zero_padded_length,), "spectrum", torch.float64)
    if out is None:
        out = torch.empty_like(proj)
    _require(out, tuple(proj.shape), "out")
    if out.data_ptr() == proj.data_ptr():
        raise ValueError("pixel_filter writes mirrored rows and cannot run in place")
    lib = _nat.lib()
    g = scan.ctypes_struct()
    nbytes = lib.ctr_fdk_filter_workspace_bytes(ctypes.byref(g))
    ws = _workspace(nbytes, dev)
    _nat.check(lib.ctr_pixel_filter(ctypes.byref(g), _ptr(spectrum), _ptr(proj), _ptr(out),
                                 int(proj.shape[0]), _ptr(ws), nbytes, _stream(dev)), "ctr_pixel_filter")
    return out


def pixel_backproject(q: torch.Tensor, scan: PixelScan, scale: float | None = None,
                      out: torch.Tensor | None = None, accumulate: bool = False,
                      angles: torch.Tensor | None = None) -> torch.Tensor:
    """Voxel-driven bilinear backprojection of Q [n_views, nv, nu] into [nz, ny, nx]
    (python/ConeBeamGPU.py:116-129, 65-111, 303-315).  ``scale`` defaults to ``dtheta``."""
    _require(q, scan.proj_shape, "q")
    dev = q.device
    if out is None:
        out = torch.empty(scan.vol_shape, dtype=torch.float32, device=dev)
        accumulate = False
    _require(out, scan.vol_shape, "vol")
    if angles is None:
        angles = scan.device_angles(dev)
    _require(angles, (scan.n_views,), "angles", torch.float64)
    lib = _nat.lib()
    g = scan.ctypes_struct()
    nbytes = lib.ctr_pixel_backproject_workspace_bytes(ctypes.byref(g))
    ws = _workspace(nbytes, dev)
    _nat.check(lib.ctr_pixel_backproject(ctypes.byref(g), _ptr(angles), _ptr(q), _ptr(out),
                                      float(scan.dtheta if scale is None else scale),
                                      int(bool(accumulate)), _ptr(ws), nbytes, _stream(dev)),
            "ctr_pixel_backproject")
    return out


def pixel_fdk(proj: torch.Tensor, scan: PixelScan, filter_type: str = DEFAULT_FILTER,
              cutoff: float = DEFAULT_CUTOFF) -> torch.Tensor:
    """Filter + backprojection of ``ConeBeam.Reconstruction`` (python/ConeBeamGPU.py:292-315)."""
    return pixel_backproject(pixel_filter(proj, scan, filter_type=filter_type, cutoff=cutoff), scan)


class ConeBeam:
    """Drop-in for ``ConeBeam`` of python/ConeBeamGPU.py:163-340.

    ``params`` is read from the geometry file as the reference reads it; ``proj`` is the
    reference's host array [DetectorHeight, DetectorWidth, NumberOfViews]; ``Reconstruction``
    writes the volume to ``savefile`` in the reference's flat [z][y][x] order and keeps it in
    ``recon`` with the reference's shape [ReconX, ReconY, ReconZ] (a reshape of that flat order).
    Unknown keys and precisions raise ``ParamsFileError`` where the reference prints and goes on."""

    def __init__(self, filename: str, device="cuda"):
        self.params = read_params_file(filename, keys=PIXEL_KEYS)
        self.params["fov"], self.params["fovz"] = 0, 0
        self.device = torch.device(device)
        p = self.params
        self.proj = np.zeros([int(p["DetectorHeight"]), int(p["DetectorWidth"]), int(p["NumberOfViews"])])
        self.recon = np.zeros([p["ReconX"], p["ReconY"], p["ReconZ"]])

    def LoadData(self):
        """One ``[DetectorHeight, DetectorWidth]`` image per ``*.dat`` file under ``DataPath``,
        in sorted order (python/ConeBeamGPU.py:322-346)."""
        p = self.params
        try:
            dtype = PIXEL_PRECISIONS[p["precision"]]
        except KeyError:
            raise ParamsFileError("Unknown data precision") from None
        files = sorted(glob.glob(p["DataPath"] + "*.dat"))
        ns, nx, ny = int(p["NumberOfViews"]), int(p["DetectorWidth"]), int(p["DetectorHeight"])
        if len(files) != ns:
            raise ParamsFileError("Number of file is different from number of projection data required")
        for c, f in enumerate(files):
            self.proj[:, :, c] = np.fromfile(f, dtype=dtype).reshape([ny, nx])

    def Reconstruction(self, savefile: str, filter_type: str = DEFAULT_FILTER,
                       cutoff: float = DEFAULT_CUTOFF):
        scan = PixelScan(self.params)
        self.params["fov"], self.params["fovz"] = scan.fov, scan.fovz
        proj = torch.from_numpy(np.ascontiguousarray(np.transpose(self.proj, (2, 0, 1)), dtype=np.float32))
        proj = proj.to(self.device)
        vol = pixel_fdk(proj, scan, filter_type, cutoff)
        flat = vol.cpu().numpy().reshape(-1)
        recon = flat.reshape([scan.nx, scan.ny, scan.nz])
        self.recon = recon.astype(np.float32)
        recon.tofile(savefile, sep="", format="")
        return self.recon

    @staticmethod
    def Filter(N: int, pixel_size: float, FilterType: str, cutoff: float, device="cuda") -> np.ndarray:
        """``ConeBeam.Filter``: the centred magnitude of length N-1, computed on the GPU."""
        h = pixel_spectrum(int(N) - 1, pixel_size, FilterType, cutoff, device)
        return np.fft.fftshift(h.cpu().numpy())
