"""CPU checks of the C ABI: the library loads, exports every declared symbol, and validates
arguments on the host (no kernel is launched on these paths)."""
import ctypes
import os
import re

import numpy as np
import pytest

from ctreconstruction_amd import _native as N
from ctreconstruction_amd.geometry import ScanGeometry
from tests.conftest import make_params

HEADER = os.path.join(N.REPO_DIR, "include", "ctrecon.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ctr_[a-z_0-9]+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("HIP library not built (run __graft_entry__.build())")
    return N.lib()


def test_header_and_binding_agree():
    assert declared_functions() == sorted(N.EXPORTS)


def test_library_exports_every_declared_symbol(lib):
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.ctr_version() == 1


def test_struct_layout_matches_header():
    # int32 x 8, then 7 doubles, 3 + 3 + 1 + 3 + 2 doubles
    assert ctypes.sizeof(N.CtrGeometry) == 8 * 4 + (7 + 3 + 3 + 1 + 3 + 2) * 8


def test_strerror(lib):
    assert lib.ctr_strerror(0) == b"ok"
    assert b"workspace" in lib.ctr_strerror(3)
    assert lib.ctr_strerror(99) == b"unknown error"


def test_zero_padded_length(lib):
    for nu in (2, 5, 256, 384, 512, 513, 1024, 4097):
        assert lib.ctr_zero_padded_length(nu) == ScanGeometry(make_params(NumberOfDetectorPixels=[nu, 4])).zero_padded_length
    assert lib.ctr_zero_padded_length(1) == -1


@pytest.mark.parametrize("n", [8, 360, 720, 45])
def test_branch_selection_matches_host_geometry(lib, n):
    geom = ScanGeometry(make_params(NumberOfViews=n))
    g = geom.ctypes_struct()
    ang = np.ascontiguousarray(geom.angles)
    out = np.zeros(n, dtype=np.int32)
    rc = lib.ctr_select_branches(ctypes.byref(g), ang.ctypes.data_as(ctypes.c_void_p),
                                 out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    np.testing.assert_array_equal(out, geom.branch)
    # circular orbit: x-slabs where |sin| dominates
    assert set(np.unique(out)) <= {0, 1}


def test_branch_selection_marks_z_slabs(lib):
    geom = ScanGeometry(make_params(SourceInit=[0, 10.0, 900.0], DetectorInit=[0, -10.0, -900.0]))
    g = geom.ctypes_struct()
    ang = np.ascontiguousarray(geom.angles)
    out = np.zeros(geom.nview, dtype=np.int32)
    rc = lib.ctr_select_branches(ctypes.byref(g), ang.ctypes.data_as(ctypes.c_void_p),
                                 out.ctypes.data_as(ctypes.c_void_p))
    # |S_z - D_z| = 1800 beats SDD*max(|sin|,|cos|) away from the axes: z-slab views, which the
    # forward projector takes (the matched adjoint does not)
    assert rc == N.CTR_OK
    assert out[1] == 2 and out[0] == 1
    np.testing.assert_array_equal(out, geom.branch)
    with pytest.raises(N.CtrError):
        geom.check_supported()


def test_invalid_geometry_is_rejected_before_launch(lib):
    geom = ScanGeometry(make_params())
    g = geom.ctypes_struct()
    g.nx = 0
    dummy = ctypes.c_void_p(16)  # never dereferenced: validation fails first
    rc = lib.ctr_dd_forward(ctypes.byref(g), dummy, dummy, dummy, dummy, dummy, 1 << 30, None)
    assert rc == N.CTR_ERR_INVALID
    g = geom.ctypes_struct()
    g.dz = -0.5
    assert lib.ctr_dd_backproject(ctypes.byref(g), dummy, dummy, dummy, dummy, 0, 1.0, 0.0, 0,
                                  dummy, 1 << 30, None) == N.CTR_ERR_INVALID
    # window rows are addressed with 24-bit multiplies by nu
    g = geom.ctypes_struct()
    g.nu, g.nv, g.n_views = 1 << 23, 1, 1
    assert lib.ctr_dd_backproject(ctypes.byref(g), dummy, dummy, dummy, dummy, 0, 1.0, 0.0, 0,
                                  dummy, 1 << 30, None) == N.CTR_ERR_INVALID
    g = geom.ctypes_struct()
    g.detector_shape = 7
    assert lib.ctr_fdk_filter(ctypes.byref(g), dummy, dummy, dummy, 1, dummy, 1 << 30, None) == \
        N.CTR_ERR_UNSUPPORTED


def test_workspace_too_small_is_rejected(lib):
    geom = ScanGeometry(make_params())
    g = geom.ctypes_struct()
    need = lib.ctr_dd_forward_workspace_bytes(ctypes.byref(g))
    # tables + two z-prefix tables of the volume
    assert need >= 2 * 16 * 16 * 16 * 4
    dummy = ctypes.c_void_p(16)
    assert lib.ctr_dd_forward(ctypes.byref(g), dummy, dummy, dummy, dummy, dummy, need - 1, None) == \
        N.CTR_ERR_WORKSPACE
    nbp = lib.ctr_dd_backproject_workspace_bytes(ctypes.byref(g), 0)
    assert lib.ctr_dd_backproject(ctypes.byref(g), dummy, dummy, dummy, dummy, 0, 1.0, 0.0, 0,
                                  dummy, nbp - 1, None) == N.CTR_ERR_WORKSPACE


def test_forward_offset_limit_is_rejected_before_launch(lib):
    # 256 (nz + 2) max(nx, ny) >= 2^31: a batch of 64 slabs would need offsets past 31 bits
    geom = ScanGeometry(make_params())
    g = geom.ctypes_struct()
    g.nx, g.ny, g.nz = 8192, 1, 1024
    need = lib.ctr_dd_forward_workspace_bytes(ctypes.byref(g))
    assert need > 0
    dummy = ctypes.c_void_p(16)  # never dereferenced: the check comes before any launch
    assert lib.ctr_dd_forward(ctypes.byref(g), dummy, dummy, dummy, dummy, dummy, need, None) == \
        N.CTR_ERR_UNSUPPORTED


def test_bad_filter_arguments_are_rejected(lib):
    geom = ScanGeometry(make_params())
    g = geom.ctypes_struct()
    dummy = ctypes.c_void_p(16)
    assert lib.ctr_filter_spectrum(ctypes.byref(g), 9, 1.0, dummy, None) == N.CTR_ERR_INVALID
    assert lib.ctr_filter_spectrum(ctypes.byref(g), 4, 0.0, dummy, None) == N.CTR_ERR_INVALID
    g.nu = 5000  # FFT length 16384 > 8192
    assert lib.ctr_filter_spectrum(ctypes.byref(g), 0, 1.0, dummy, None) == N.CTR_ERR_UNSUPPORTED


def test_product_path_fails_loudly_without_gpu_tensors():
    import torch

    from ctreconstruction_amd import projector as P
    geom = ScanGeometry(make_params())
    with pytest.raises(ValueError, match="GPU"):
        P.dd_forward(torch.zeros(geom.vol_shape), geom)


def _pixel_struct(**kw):
    from ctreconstruction_amd.pixel_fdk import PixelScan
    from tests.conftest import make_cone_params
    return PixelScan(make_cone_params(**kw)).ctypes_struct()


def test_pixel_backproject_rejects_unsupported_geometry_before_launch(lib):
    dummy = ctypes.c_void_p(16)  # never dereferenced: validation fails first

    def bp(g, ws=1 << 30):
        return lib.ctr_pixel_backproject(ctypes.byref(g), dummy, dummy, dummy, 1.0, 0, dummy, ws, None)

    g = _pixel_struct()
    assert lib.ctr_pixel_backproject_workspace_bytes(ctypes.byref(g)) > 0
    g.detector_shape = N.CTR_DETECTOR_CURVED
    assert bp(g) == N.CTR_ERR_UNSUPPORTED
    g = _pixel_struct()
    g.det_offset[0] = 0.5
    assert bp(g) == N.CTR_ERR_UNSUPPORTED
    g = _pixel_struct()
    g.nv = 1  # the reference indexes p[1]
    assert bp(g) == N.CTR_ERR_INVALID
    assert lib.ctr_pixel_backproject_workspace_bytes(ctypes.byref(g)) == 0
    g = _pixel_struct()
    g.sad = 10.0  # source inside the voxel grid
    assert bp(g) == N.CTR_ERR_UNSUPPORTED
    g = _pixel_struct()
    need = lib.ctr_pixel_backproject_workspace_bytes(ctypes.byref(g))
    assert bp(g, need - 1) == N.CTR_ERR_WORKSPACE
    assert lib.ctr_pixel_backproject(ctypes.byref(g), dummy, dummy, dummy, float("nan"), 0, dummy,
                                     need, None) == N.CTR_ERR_INVALID


def test_pixel_filter_arguments_are_rejected_before_launch(lib):
    dummy = ctypes.c_void_p(16)
    assert lib.ctr_pixel_filter_spectrum(48, 0.5, 0, 0.3, dummy, None) == N.CTR_ERR_INVALID  # not 2^k
    assert lib.ctr_pixel_filter_spectrum(16384, 0.5, 0, 0.3, dummy, None) == N.CTR_ERR_UNSUPPORTED
    assert lib.ctr_pixel_filter_spectrum(64, 0.0, 0, 0.3, dummy, None) == N.CTR_ERR_INVALID
    assert lib.ctr_pixel_filter_spectrum(64, 0.5, 4, 0.0, dummy, None) == N.CTR_ERR_INVALID
    g = _pixel_struct()
    # in place is refused: rows are written mirrored
    assert lib.ctr_pixel_filter(ctypes.byref(g), dummy, dummy, dummy, 1, dummy, 1 << 30, None) == \
        N.CTR_ERR_INVALID


def test_pixel_scan_places_voxels_on_the_reference_linspace():
    from ctreconstruction_amd.pixel_fdk import PixelScan
    from oracle import pixel_oracle as PO
    from tests.conftest import make_cone_params
    for kw in ({}, {"ReconX": 7, "ReconY": 1, "ReconZ": 5}):
        prm = make_cone_params(**kw)
        g = PixelScan(prm).ctypes_struct()
        ref = PO.PixelGeometry(prm)
        for n, d, c, axis in ((g.nx, g.dx, g.recon_center[0], ref.x), (g.ny, g.dy, g.recon_center[1], ref.y),
                              (g.nz, g.dz, g.recon_center[2], ref.z)):
            np.testing.assert_allclose((np.arange(n) - (n - 1) / 2) * d + c, axis, rtol=0, atol=1e-12)
