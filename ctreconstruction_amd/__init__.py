"""MI355X-native distance-driven CT projector pair and FDK (drop-in for leehoy/CTReconstruction).

Host code on PyTorch-ROCm; compute in hand-written gfx950 HIP kernels behind the C ABI of
``include/ctrecon.h`` (library ``lib/libctrecon.so``).
"""
from .geometry import ScanGeometry, merged_params  # noqa: F401
from . import projector  # noqa: F401
from .projector import dd_backproject, dd_forward, fdk, fdk_filter, filter_spectrum  # noqa: F401
from .reconstruction import Backward, Forward, Reconstruction  # noqa: F401

__all__ = ["ScanGeometry", "Reconstruction", "Forward", "Backward", "dd_forward", "fdk_filter",
           "filter_spectrum", "dd_backproject", "fdk", "projector", "merged_params"]
