import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


def make_params(**kw):
    """Reference-style params dict (python/FDK_512.py:13-17 layout) for small test geometries."""
    p = {
        "SourceInit": [0, 1000.0, 0], "DetectorInit": [0, -500.0, 0], "StartAngle": 0,
        "EndAngle": 2 * np.pi, "NumberOfDetectorPixels": [21, 15], "DetectorPixelSize": [1.0, 1.0],
        "NumberOfViews": 8, "ImagePixelSpacing": [0.5, 0.5, 0.5], "NumberOfImage": [16, 16, 16],
        "PhantomCenter": [0, 0, 0], "RotationOrigin": [0, 0, 0], "ReconCenter": [0, 0, 0],
        "Method": "Distance", "FilterType": "ram-lak", "cutoff": 1, "GPU": 1,
        "DetectorShape": "Flat", "Pitch": 0, "DetectorOffset": [0, 0],
    }
    p.update(kw)
    return p


def make_cone_params(**kw):
    """``ConeBeam`` params dict (python/ConeBeamGPU.py:164-197; AngleCoverage in radians)."""
    p = {
        "SourceToAxis": 100.0, "SourceToDetector": 150.0, "DetectorPixelWidth": 0.8,
        "DetectorPixelHeight": 0.8, "DetectorWidth": 48, "DetectorHeight": 40,
        "NumberOfViews": 48, "AngleCoverage": 2 * np.pi, "ReconX": 24, "ReconY": 24,
        "ReconZ": 24, "DataPath": "", "precision": "float32",
    }
    p.update(kw)
    return p


@pytest.fixture
def params_factory():
    return make_params
