"""The reference's ``key: value`` geometry text file and its mapping onto the params dict.

Format (python/ReconstructionParams.txt:1-11, python/ReconstructionParamsFan.txt:1-10), parsed
as ``ConeBeam.__init__`` / ``FanBeam.__init__`` parse it (python/ConeBeam.py:51-83,
python/FanBeam.py:43-75):

* one ``Key: value`` per line, the key is the text before the first ``:``, stripped;
* ``AngleCoverage`` is in degrees and is stored in radians;
* ``ReconVolume: X*Y*Z`` (cone) or ``X*Y`` (fan) sets ``ReconX``, ``ReconY`` (, ``ReconZ``);
* ``DataPath`` and ``precision`` stay strings, every other known key is a float;
* any other key is an error ("Unknown variables", ``ErrorDescription(1)``).

The reference prints that error and keeps the keys read before it; here it raises
``ParamsFileError`` with the same message and the offending line, so that a half-read geometry
never reaches the projectors.

The file gives no voxel size.  ``ConeBeam.Reconstruction`` places its voxel centres at
``linspace(-fov/2, fov/2, ReconX)`` with ``fov = 2 SAD sin(atan(W du / 2 / SDD))`` and likewise
``fovz`` from the detector height (python/ConeBeam.py:103-112).  ``to_recon_params`` reproduces
exactly those centres on the params-dict grid (centres at ``(i - (n-1)/2) * spacing``) by taking
``spacing = fov / (ReconX - 1)``.
"""
from __future__ import annotations

import math

import numpy as np

CONE_KEYS = ("SourceToDetector", "SourceToAxis", "DataPath", "precision", "AngleCoverage",
             "ReconX", "ReconY", "ReconZ", "DetectorPixelHeight", "DetectorPixelWidth",
             "DetectorWidth", "DetectorHeight", "NumberOfViews", "fov", "fovz")
FAN_KEYS = ("SourceToDetector", "SourceToAxis", "DataPath", "precision", "AngleCoverage",
            "ReconX", "ReconY", "DetectorPixelWidth", "DetectorWidth", "DetectorHeight",
            "NumberOfViews", "fov", "fovz")
PRECISIONS = {"float32": np.float32, "float64": np.float64}


class ParamsFileError(ValueError):
    pass


def read_params_file(filename: str, fan: bool = False, keys: tuple | None = None) -> dict:
    """Parse a geometry text file into the ``ConeBeam`` (or ``FanBeam``) params dict.

    Every key starts at the reference's default (0 or ''); ``AngleCoverage`` is returned in
    radians, as the reference stores it.  ``keys`` replaces the accepted key set (the GPU
    ``ConeBeam`` of python/ConeBeamGPU.py:164-168 also accepts ``Mode``)."""
    if keys is None:
        keys = FAN_KEYS if fan else CONE_KEYS
    params = {k: ("" if k in ("DataPath", "precision") else 0) for k in keys}
    with open(filename) as f:
        for lineno, line in enumerate(f, 1):
            if not line.strip():
                continue  # blank lines (e.g. a trailing one) carry no key
            key = line.split(":")[0].strip()
            if key not in params and key != "ReconVolume":
                raise ParamsFileError(f"Unknown variables: {filename}:{lineno}: {line.rstrip()!r}")
            value = line.split(":")[1].strip()
            if key == "AngleCoverage":
                params[key] = float(value) * math.pi / 180.0
            elif key == "ReconVolume":
                dims = value.split("*")
                need = 2 if fan else 3
                if len(dims) < need:
                    raise ParamsFileError(f"ReconVolume needs {need} sizes: {filename}:{lineno}")
                params["ReconX"], params["ReconY"] = int(dims[0]), int(dims[1])
                if not fan:
                    params["ReconZ"] = int(dims[2])
            elif key in ("DataPath", "precision"):
                params[key] = value
            else:
                params[key] = float(value)
    return params


def field_of_view(params: dict) -> tuple[float, float]:
    """(fov, fovz) as ``ConeBeam.Reconstruction`` derives them (python/ConeBeam.py:103-107)."""
    sad = float(params["SourceToAxis"])
    sdd = float(params["SourceToDetector"])
    width = params["DetectorWidth"] * params["DetectorPixelWidth"]
    height = params["DetectorHeight"] * params.get("DetectorPixelHeight", 0.0)
    fov = 2.0 * sad * math.sin(math.atan(width / 2.0 / sdd))
    fovz = 2.0 * sad * math.sin(math.atan(height / 2.0 / sdd))
    return fov, fovz


def data_dtype(params: dict):
    """numpy dtype of the ``precision`` entry ("Unknown data precision", ``ErrorDescription(2)``)."""
    try:
        return PRECISIONS[params["precision"]]
    except KeyError:
        raise ParamsFileError(f"Unknown data precision: {params['precision']!r}") from None


def to_recon_params(file_params: dict, **overrides) -> dict:
    """Map a parsed geometry file onto the ``Reconstruction`` params dict.

    Source on +y at SAD, detector on -y at SDD - SAD (the scripts' convention,
    python/FDK_512.py:13-17); angles ``linspace(0, AngleCoverage, n+1)[:-1]`` as in
    python/ConeBeam.py:111-112; voxel spacing ``fov / (n - 1)`` so that voxel centres equal
    the reference's ``linspace(-fov/2, fov/2, n)``.  A fan file (no ``ReconZ``) maps to one
    detector row and one slice, with the slice and row pitch set to the magnified and detector
    column pitch.  ``overrides`` replace entries of the result (e.g. ``FilterType``)."""
    sad = float(file_params["SourceToAxis"])
    sdd = float(file_params["SourceToDetector"])
    nx, ny = int(file_params["ReconX"]), int(file_params["ReconY"])
    fov, fovz = field_of_view(file_params)
    du = float(file_params["DetectorPixelWidth"])
    if "ReconZ" in file_params and int(file_params["ReconZ"]) > 0:
        nz = int(file_params["ReconZ"])
        nu, nv = int(file_params["DetectorWidth"]), int(file_params["DetectorHeight"])
        dv = float(file_params["DetectorPixelHeight"])
        dz = fovz / (nz - 1) if nz > 1 else dv * sad / sdd
    else:
        nz, nu, nv = 1, int(file_params["DetectorWidth"]), 1
        dv = du
        dz = du * sad / sdd
    spacing = lambda n: fov / (n - 1) if n > 1 else du * sad / sdd  # noqa: E731
    params = {
        "SourceInit": [0.0, sad, 0.0], "DetectorInit": [0.0, -(sdd - sad), 0.0],
        "StartAngle": 0.0, "EndAngle": float(file_params["AngleCoverage"]),
        "NumberOfDetectorPixels": [nu, nv], "DetectorPixelSize": [du, dv],
        "NumberOfViews": int(file_params["NumberOfViews"]),
        "ImagePixelSpacing": [spacing(nx), spacing(ny), dz], "NumberOfImage": [nx, ny, nz],
        "PhantomCenter": [0, 0, 0], "RotationOrigin": [0, 0, 0], "ReconCenter": [0, 0, 0],
        "Method": "Distance", "FilterType": "ram-lak", "cutoff": 1, "GPU": 1,
        "DetectorShape": "Flat", "Pitch": 0, "DetectorOffset": [0, 0],
    }
    params.update(overrides)
    return params


def load_recon_params(filename: str, fan: bool | None = None, **overrides) -> dict:
    """``read_params_file`` + ``to_recon_params``; ``fan=None`` picks fan for a 2-size ReconVolume."""
    if fan is None:
        with open(filename) as f:
            vol = [ln.split(":")[1] for ln in f if ln.split(":")[0].strip() == "ReconVolume"]
        fan = bool(vol) and len(vol[-1].split("*")) == 2
    return to_recon_params(read_params_file(filename, fan=fan), **overrides)
