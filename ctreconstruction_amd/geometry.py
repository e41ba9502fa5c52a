"""Scan geometry: the reference's params dict, its derived scalars, and their device form.

``ScanGeometry`` derives what ``Reconstruction.__init__`` derives
(python/Reconstruction_pycuda.py:64-99): SAD, SDD, helical advance, the angle grid
``linspace(StartAngle, EndAngle, n+1)[:-1]`` (:788-789) and ``dtheta`` (:349).  It also fixes
the slab branch of every view on the host with the reference's own FP64 comparisons
(:828-830, :852-853, :921-922), so that ties at 45-degree views break exactly as they do in
the reference, and uploads angles and branches once per device.  Everything else (detector
lattice, footprint tables) is computed on the GPU in FP64 by the library.
"""
from __future__ import annotations

import math

import numpy as np

from . import _native as N

# Defaults of the reference's params dict (python/Reconstruction_pycuda.py:65-70).
DEFAULT_PARAMS = {
    "SourceInit": [0, 0, 0], "DetectorInit": [0, 0, 0], "StartAngle": 0, "EndAngle": 0,
    "NumberOfDetectorPixels": [0, 0], "DetectorPixelSize": [0, 0], "NumberOfViews": 0,
    "ImagePixelSpacing": [0, 0, 0], "NumberOfImage": [0, 0, 0], "PhantomCenter": [0, 0, 0],
    "RotationOrigin": [0, 0, 0], "Method": "Distance", "FilterType": "ram-lak",
    "ReconCenter": [0, 0, 0], "cutoff": 1, "GPU": 0, "DetectorShape": "Flat", "Pitch": 0,
    "DetectorOffset": [0, 0],
}


def merged_params(params: dict) -> dict:
    """The reference replaces its defaults wholesale (``self.params = params``, :71) and then
    fails on a missing key; here missing keys take the defaults instead.  The legacy scripts'
    ``Origin`` key (python/SART.py:14) stands in for ``RotationOrigin`` when that is absent."""
    p = dict(DEFAULT_PARAMS)
    if "RotationOrigin" not in params and "Origin" in params:
        p["RotationOrigin"] = params["Origin"]
    p.update(params)
    return p


class ScanGeometry:
    """Derived geometry of one circular / helical cone-beam (or fan-beam, nz = nv = 1) scan."""

    def __init__(self, params: dict):
        p = merged_params(params)
        self.params = p
        self.nu, self.nv = (int(v) for v in p["NumberOfDetectorPixels"])
        self.du, self.dv = (float(v) for v in p["DetectorPixelSize"])
        self.nx, self.ny, self.nz = (int(v) for v in p["NumberOfImage"])
        self.dx, self.dy, self.dz = (float(v) for v in p["ImagePixelSpacing"])
        self.origin = np.asarray(p["RotationOrigin"], dtype=np.float64)
        self.source = np.asarray(p["SourceInit"], dtype=np.float64)
        self.detector = np.asarray(p["DetectorInit"], dtype=np.float64)
        self.sad = float(np.sqrt(np.sum((self.source - self.origin) ** 2.0)))
        self.sdd = float(np.sqrt(np.sum((self.source - self.detector) ** 2.0)))
        self.pitch = float(p["Pitch"])
        self.helical = self.pitch * (self.nv * self.dv) * self.sad / self.sdd
        self.nview = int(p["NumberOfViews"])
        self.start_angle = float(p["StartAngle"])
        self.end_angle = float(p["EndAngle"])
        self.angles = np.linspace(self.start_angle, self.end_angle, self.nview + 1)[:-1]
        self.dtheta = (float(self.angles[1] - self.angles[0]) if self.nview > 1
                       else self.end_angle - self.start_angle)
        self.recon_center = np.asarray(p["ReconCenter"], dtype=np.float64)
        self.det_offset = np.asarray(p.get("DetectorOffset", [0, 0]), dtype=np.float64)
        self.shape = p["DetectorShape"]
        if self.shape not in ("Flat", "Curved"):
            raise ValueError(f"DetectorShape must be 'Flat' or 'Curved', got {self.shape!r}")
        self.filter_type = p["FilterType"]
        self.cutoff = float(p["cutoff"])
        self.branch = self._branches()
        self._device_cache = {}

    # -- per-view slab branch, evaluated exactly as the reference's host loop does
    def _branches(self) -> np.ndarray:
        H, sad, sdd = self.helical, self.sad, self.sdd
        zs, zd = self.source[2], self.detector[2]
        out = np.empty(self.nview, dtype=np.int32)
        for i, th in enumerate(self.angles):
            S = np.array([-sad * np.sin(th), sad * np.cos(th), zs + H * th / (2 * np.pi)])
            D = np.array([(sdd - sad) * np.sin(th), -(sdd - sad) * np.cos(th), zd + H * th / (2 * np.pi)])
            ax, ay, az = abs(S[0] - D[0]), abs(S[1] - D[1]), abs(S[2] - D[2])
            if ax >= ay and ax >= az:
                out[i] = N.CTR_BRANCH_X
            elif ay >= ax and ay >= az:
                out[i] = N.CTR_BRANCH_Y
            else:
                out[i] = N.CTR_BRANCH_Z
        return out

    @property
    def zero_padded_length(self) -> int:
        """``2**ceil(log2(2*(nu-1)))`` (python/Reconstruction_pycuda.py:236)."""
        return int(2 ** (math.ceil(math.log2(2.0 * (self.nu - 1)))))

    @property
    def tau(self) -> float:
        """Detector sampling at the isocentre, du*SAD/SDD (python/Reconstruction_pycuda.py:242)."""
        return self.du * self.sad / self.sdd

    @property
    def vol_shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def proj_shape(self):
        return (self.nview, self.nv, self.nu)

    def ctypes_struct(self, nview: int | None = None) -> N.CtrGeometry:
        g = N.CtrGeometry()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.nu, g.nv = self.nu, self.nv
        g.n_views = self.nview if nview is None else nview
        g.detector_shape = N.CTR_DETECTOR_FLAT if self.shape == "Flat" else N.CTR_DETECTOR_CURVED
        g.dx, g.dy, g.dz = self.dx, self.dy, self.dz
        g.du, g.dv = self.du, self.dv
        g.sad, g.sdd = self.sad, self.sdd
        g.source_init[:] = [float(v) for v in self.source]
        g.detector_init[:] = [float(v) for v in self.detector]
        g.helical = self.helical
        g.recon_center[:] = [float(v) for v in self.recon_center]
        g.det_offset[:] = [float(v) for v in self.det_offset]
        return g

    def device_views(self, device):
        """(angles float64, branch int32) device tensors, uploaded once per device."""
        import torch
        key = str(device)
        if key not in self._device_cache:
            ang = torch.from_numpy(np.ascontiguousarray(self.angles, dtype=np.float64)).to(device)
            br = torch.from_numpy(np.ascontiguousarray(self.branch, dtype=np.int32)).to(device)
            self._device_cache[key] = (ang, br)
        return self._device_cache[key]

    def with_views(self, views) -> "ScanGeometry":
        """The same scan restricted to ``views`` (their angles and branches; dtheta of the full scan)."""
        import copy
        sub = copy.copy(self)
        idx = np.asarray(views, dtype=np.int64)
        sub.angles = self.angles[idx].copy()
        sub.branch = self.branch[idx].copy()
        sub.nview = len(idx)
        sub._device_cache = {}
        return sub

    def check_supported(self):
        """The matched adjoint (and so autograd, SART, ASD-POCS) covers x- and y-slab views only;
        the forward projector and the FDK backprojector take every view."""
        if np.any(self.branch == N.CTR_BRANCH_Z):
            raise N.CtrError(N.CTR_ERR_UNSUPPORTED,
                             "matched adjoint of z-slab views (|S_z - D_z| dominates) is not built")
