"""Drop-in ``Reconstruction`` class: the reference's host operator API on the HIP kernels.

Mirrors ``Reconstruction`` of python/Reconstruction_pycuda.py:62-1059 (and its numba twin
python/Reconstruction_numba.py:55-1117): same constructor argument (the params dict), same
attributes (``image [nz,ny,nx]``, ``proj [nView,nv,nu]``, ``params``, ``nu``/``nv``/...,
``SAD``/``SDD``/``HelicalTrans``), same methods (``forward``, ``Filtering``, ``backward``,
``LoadProj``/``LoadRecon``/``SaveProj``/``SaveRecon``, static ``Filter``,
``Flat/CurvedDetectorConstruction``, ``FlipProj``) and the same raw little-endian float32 file
layout.  ``image``/``proj`` may be NumPy arrays (results come back as NumPy, as in the
reference) or torch tensors on the GPU (results stay on the GPU, no host round trip).
The ``GPU`` key is accepted and ignored: every operator runs on the GPU.
"""
from __future__ import annotations

import warnings

import numpy as np
import torch

from . import projector as P
from .geometry import ScanGeometry, merged_params


VARIANTS = ("pycuda", "numba")


class Reconstruction:
    """``variant`` picks which of the reference's two GPU classes to follow where they differ:
    ``"pycuda"`` (python/Reconstruction_pycuda.py, the default) builds the forward projector's
    voxel planes on ``ReconCenter`` (:792-794) and backprojects without weight zeroing;
    ``"numba"`` (python/Reconstruction_numba.py) builds them on ``PhantomCenter`` (:835-837) and
    zeroes backprojection weights below 1e-5 (python/GPUFuncs_numba.py:28-29, 45-46).  Both
    backproject onto ``ReconCenter``."""

    def __init__(self, params: dict, device=None, variant: str = "pycuda"):
        if variant not in VARIANTS:
            raise ValueError(f"variant must be one of {VARIANTS}, got {variant!r}")
        self.variant = variant
        self.geom = ScanGeometry(params)
        g = self.geom
        if variant == "numba":
            self._fp_geom = ScanGeometry(dict(g.params, ReconCenter=[float(v) for v in g.params["PhantomCenter"]]))
        else:
            self._fp_geom = g
        self.params = g.params
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        # attribute names of python/Reconstruction_pycuda.py:72-99
        self.nu, self.nv = g.nu, g.nv
        self.du, self.dv = g.du, g.dv
        self.nx, self.ny, self.nz = g.nx, g.ny, g.nz
        self.dx, self.dy, self.dz = g.dx, g.dy, g.dz
        self.Origin = g.origin
        self.PhantomCenter = np.array(self.params["PhantomCenter"])
        self.cutoff = g.cutoff
        self.DetectorShape = g.shape
        self.P = g.pitch
        self.Source = g.source.copy()
        self.Detector = g.detector.copy()
        self.SAD = g.sad
        self.SDD = g.sdd
        self.HelicalTrans = g.helical
        self.nView = g.nview
        self.sAngle = g.start_angle
        self.eAngle = g.end_angle
        span = self.eAngle - self.sAngle
        self.Proj2pi = self.nView / (span / (2 * np.pi)) if span != 0 else float("inf")
        self.Method = self.params["Method"]
        self.FilterType = g.filter_type
        self.source_z0 = g.source[2]
        self.detector_z0 = g.detector[2]
        self.ReconCenter = self.params["ReconCenter"]
        self.DetectorOffset = self.params["DetectorOffset"]
        self.GPU = True
        self.image = None
        self.proj = None
        self._spectrum = None

    @classmethod
    def from_params_file(cls, filename, device=None, **overrides) -> "Reconstruction":
        """Geometry from a ``key: value`` text file (python/ReconstructionParams.txt:1-11, parsed as
        python/ConeBeam.py:51-83); voxel centres match python/ConeBeam.py:103-112."""
        from .paramsfile import load_recon_params
        return cls(load_recon_params(filename, **overrides), device=device)

    # ------------------------------------------------------------------ raw file I/O (:101-111)
    def LoadProj(self, filename, dtype=np.float32):
        self.proj = np.fromfile(filename, dtype=dtype).reshape([self.nView, self.nv, self.nu])

    def LoadRecon(self, filename, dtype=np.float32):
        self.image = np.fromfile(filename, dtype=dtype).reshape([self.nz, self.ny, self.nx])

    def SaveProj(self, filename):
        _host(self.proj).tofile(filename, sep="", format="")

    def SaveRecon(self, filename):
        _host(self.image).tofile(filename, sep="", format="")

    # ------------------------------------------------------------------ operators
    def _to_device(self, a, shape, name):
        if a is None:
            raise ValueError(f"{name} is not set")
        if isinstance(a, torch.Tensor):
            t = a.to(device=self.device, dtype=torch.float32)
        else:
            t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(self.device)
        return t.reshape(shape).contiguous()

    def _method_check(self):
        if self.Method != "Distance":
            raise NotImplementedError(f"Method {self.Method!r}: only 'Distance' is implemented "
                                      "(the reference's 'Ray' branch is an empty stub)")

    def forward(self):
        """Distance-driven forward projection of ``image`` into ``proj`` (:750-760)."""
        self._method_check()
        vol = self._to_device(self.image, self.geom.vol_shape, "image")
        proj = P.dd_forward(vol, self._fp_geom)
        self.proj = proj if isinstance(self.image, torch.Tensor) else _host(proj)

    def Filtering(self):
        """Cosine weight + ramp-family filter of ``proj`` in place (:223-229)."""
        was_tensor = isinstance(self.proj, torch.Tensor)
        proj = self._to_device(self.proj, (-1, self.nv, self.nu), "proj")
        if self._spectrum is None or self._spectrum.device != proj.device:
            self._spectrum = P.filter_spectrum(self.geom, device=proj.device)
        q = P.fdk_filter(proj, self.geom, self._spectrum, out=proj)
        self.proj = q if was_tensor else _host(q)

    def backward(self):
        """FDK-weighted DD backprojection of ``proj`` (times dtheta) into ``image`` (:310-318)."""
        self._method_check()
        proj = self._to_device(self.proj, self.geom.proj_shape, "proj")
        vol = P.dd_backproject(proj, self.geom, mode="ref", zero_tol=1e-5 if self.variant == "numba" else 0.0)
        self.image = vol if isinstance(self.proj, torch.Tensor) else _host(vol)

    # ------------------------------------------------------------------ helpers of the reference
    @staticmethod
    def Filter(N, pixel_size, FilterType, cutoff):
        """Centred filter magnitude of length N-1 (python/Reconstruction_pycuda.py:186-221),
        computed on the GPU in FP64.  N-1 must be a power of two."""
        if cutoff > 1 or cutoff < 0:
            # the reference prints this message and carries on
            warnings.warn("Cutoff have to be pose between 0 and 0.5")
        zp = int(N) - 1
        nu = zp // 2 + 1  # smallest detector with this zero-padded length
        g = ScanGeometry({"NumberOfDetectorPixels": [nu, 1], "DetectorPixelSize": [pixel_size, 1.0],
                          "SourceInit": [0, 1.0, 0], "DetectorInit": [0, 0.0, 0], "NumberOfViews": 1,
                          "NumberOfImage": [1, 1, 1], "ImagePixelSpacing": [1.0, 1.0, 1.0],
                          "FilterType": FilterType, "cutoff": cutoff})
        if g.zero_padded_length != zp:
            raise ValueError(f"N-1 = {zp} is not a zero-padded FFT length (power of two >= 2)")
        spec = P.filter_spectrum(g, FilterType, cutoff).cpu().numpy()
        return np.fft.fftshift(spec)

    def FlatDetectorConstruction(self, Source, DetectorCenter, SDD, angle):
        """(DetectorIndex [3,nv,nu], DetectorBoundary [3,nv+1,nu+1]) float32 (:139-168)."""
        return _lattice(self, Source, SDD, angle, curved=False)

    def CurvedDetectorConstruction(self, Source, DetectorCenter, SDD, angle):
        """Curved (equiangular) lattice (:113-137), including its ``- dv/2`` edge rows."""
        return _lattice(self, Source, SDD, angle, curved=True)

    def FlipProj(self, dir):
        """Flip every projection along v (dir=0) or u (dir=1) (python/Reconstruction_numba.py:106-119)."""
        if dir not in (0, 1):
            warnings.warn("Cutoff have to be pose between 0 and 0.5")  # the reference's message for code 4
            return
        axis = 1 if dir == 0 else 2
        if isinstance(self.proj, torch.Tensor):
            self.proj = torch.flip(self.proj, dims=(axis,)).contiguous()
        else:
            self.proj = np.ascontiguousarray(np.flip(self.proj, axis=axis))

    @staticmethod
    def _optimalGrid(GridSize):
        """Divisor pair near sqrt(GridSize) (:170-184); kept for API compatibility only."""
        r = int(np.sqrt(GridSize))
        if r * r == GridSize:
            return (r, r)
        divs = [d for d in range(1, GridSize + 1) if GridSize % d == 0]
        i = len(divs) // 2
        return (int(divs[i]), int(divs[i - 1]))


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return a


def _lattice(r: Reconstruction, Source, SDD, angle, curved: bool):
    eu = np.array([np.cos(angle), np.sin(angle), 0.0])
    ew = np.array([np.sin(angle), -np.cos(angle), 0.0])
    off = np.asarray(r.DetectorOffset, dtype=np.float64)
    step = r.du / r.SDD if curved else r.du
    u = (np.arange(r.nu) - (r.nu - 1) / 2.0) * step + off[0]
    v = (np.arange(r.nv) - (r.nv - 1) / 2.0) * -1.0 * r.dv + off[1]
    u2 = (np.arange(r.nu + 1) - (r.nu - 1) / 2.0) * step - step / 2.0 + off[0]
    half = -r.dv / 2.0 if curved else r.dv / 2.0
    v2 = (np.arange(r.nv + 1) - (r.nv - 1) / 2.0) * -1.0 * r.dv + half + off[1]
    out = []
    for uu, vv in ((u, v), (u2, v2)):
        U, V = np.meshgrid(uu, vv)
        arr = np.zeros((3,) + U.shape, dtype=np.float32)
        for k in range(3):
            if curved:
                arr[k] = Source[k] + SDD * np.sin(U) * eu[k] + SDD * np.cos(U) * ew[k] + V * (k == 2)
            else:
                arr[k] = Source[k] + U * eu[k] + SDD * ew[k] + V * (k == 2)
        out.append(arr)
    return out[0], out[1]


# ---------------------------------------------------------------------------------------------
# Legacy entry points: Forward(filename, params).forward() and Backward(filename, params).backward()
# ---------------------------------------------------------------------------------------------

def _legacy_params(params: dict) -> dict:
    """The legacy classes' params (python/Forward.py:443-450, python/Backward.py:692-700) on the
    ``Reconstruction`` grid.

    Their voxel edge planes are ``(PhantomCenter - n/2 + k) * d - d/2`` with d = (dx, -dy, -dz)
    (python/Forward.py:538-543) and their backprojection samples voxels at
    ``(PhantomCenter - n/2 + i) * d`` (python/Backward.py:835-840), i.e. the ``Reconstruction``
    grid with ``ReconCenter = (PC_x dx - dx/2, -PC_y dy + dy/2, -PC_z dz + dz/2)``.  ``Origin``
    gives the rotation centre (SAD), the source and detector sit at z = 0."""
    p = merged_params(params)
    dx, dy, dz = (float(v) for v in p["ImagePixelSpacing"])
    pc = [float(v) for v in p.get("PhantomCenter", [0, 0, 0])]
    p["ReconCenter"] = [pc[0] * dx - dx / 2, -pc[1] * dy + dy / 2, -pc[2] * dz + dz / 2]
    p["SourceInit"] = [float(p["SourceInit"][0]), float(p["SourceInit"][1]), 0.0]
    p["DetectorInit"] = [float(p["DetectorInit"][0]), float(p["DetectorInit"][1]), 0.0]
    p["Pitch"] = 0
    p["DetectorOffset"] = [0, 0]
    return p


class Forward:
    """``Forward(filename, params)`` loads the raw float32 volume ``[nz, ny, nx]`` from
    ``filename``; ``forward()`` sets ``proj [nViews, nv, nu]`` (python/Forward.py:441-571).

    The projection is the distance-driven projector of ``Reconstruction.forward`` on the legacy
    voxel grid (see ``_legacy_params``); the numerics are those of the main path, not of the
    legacy module's own kernels."""

    def __init__(self, filename, params, device=None):
        self.params = params
        self._r = Reconstruction(_legacy_params(params), device=device)
        self.image = self.LoadFile(filename, [self._r.nz, self._r.ny, self._r.nx])
        self.proj = None

    @staticmethod
    def LoadFile(filename, image_size, dtype=np.float32):
        return np.fromfile(filename, dtype=dtype).reshape(image_size)

    def SaveProj(self, filename):
        _host(self.proj).tofile(filename, sep="", format="")

    def forward(self):
        self._r.image = self.image
        self._r.forward()
        self.proj = self._r.proj


class Backward:
    """``Backward(filename, params)`` loads raw float32 projections ``[nViews, nv, nu]``;
    ``backward()`` filters every view (``FilterType``, ``cutoff``) and backprojects it, times
    dtheta, on true detector cell edges (python/Backward.py:808-872), setting ``recon
    [nz, ny, nx]``.  Runs the main path's filter and FDK-weighted DD backprojector (mode
    "edge") on the legacy voxel grid (see ``_legacy_params``)."""

    def __init__(self, filename, params, device=None):
        self.params = params
        self._r = Reconstruction(_legacy_params(params), device=device)
        self.proj = self.LoadProj(filename, [self._r.nView, self._r.nv, self._r.nu])
        self.recon = None

    @staticmethod
    def LoadProj(filename, image_size, dtype=np.float32):
        return np.fromfile(filename, dtype=dtype).reshape(image_size)

    @staticmethod
    def LoadFile(filename, image_size, dtype=np.float32):
        return np.fromfile(filename, dtype=dtype).reshape(image_size)

    def SaveProj(self, filename):
        _host(self.proj).tofile(filename, sep="", format="")

    def SaveRecon(self, filename):
        _host(self.recon).tofile(filename, sep="", format="")

    @staticmethod
    def Filter(N, pixel_size, FilterType, cutoff):
        return Reconstruction.Filter(N, pixel_size, FilterType, cutoff)

    def backward(self):
        r = self._r
        r._method_check()
        proj = r._to_device(self.proj, r.geom.proj_shape, "proj")
        q = P.fdk_filter(proj, r.geom)
        vol = P.dd_backproject(q, r.geom, mode="edge")
        self.recon = vol if isinstance(self.proj, torch.Tensor) else _host(vol)
