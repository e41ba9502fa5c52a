"""Parity at the full benchmark sizes, through size-independent properties and a bounded oracle
sample: the cone-beam DD path at 512^3 / 720 views / 1024x768 (0.25 mm voxels, 0.4 mm cells,
Shepp-Logan) and the pixel-driven FDK on the geometry of python/ReconstructionParams.txt
(512^3, 680 views, 1024x768 at 0.390625 mm).

The oracle sample: projections that are zero except for two views, run through the full-size
GPU filter and backprojection; the oracle filters those two views and backprojects them into
the central 8 slices (zero views add exactly nothing on the GPU).  The whole-volume sample: eight
views of a phantom scan against the oracle's C restatement (oracle/c/dd_backproject.c, the
NumPy oracle's loop in C, equal to it bit for bit), and likewise for the pixel-driven FDK
against oracle/c/pixel_backproject.c.  Bar: 1e-4 of the oracle's peak magnitude, as for the
small-size parity tests.
"""
import copy
import os

import numpy as np
import pytest

from oracle import dd_oracle as O
from oracle import pixel_oracle as PO
from tests.conftest import make_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-4
SLAB = 8

DD_FULL = make_params(NumberOfImage=[512, 512, 512], ImagePixelSpacing=[0.25, 0.25, 0.25],
                      NumberOfDetectorPixels=[1024, 768], DetectorPixelSize=[0.4, 0.4],
                      NumberOfViews=720, FilterType="shepp-logan")

# python/ReconstructionParams.txt:1-11 (AngleCoverage in radians, as the reader stores it)
PIXEL_FULL = dict(SourceToAxis=1000.0, SourceToDetector=1500.0, DetectorPixelWidth=0.390625,
                  DetectorPixelHeight=0.390625, DetectorWidth=1024, DetectorHeight=768,
                  NumberOfViews=680, AngleCoverage=2 * np.pi, ReconX=512, ReconY=512, ReconZ=512,
                  DataPath="", precision="float32")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from ctreconstruction_amd import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.fail("HIP library not built: the GPU path has no fallback")
    return torch.device("cuda:0")


def _two_view_projections(shape, views, dev, seed):
    rng = np.random.default_rng(seed)
    sample = rng.uniform(0.0, 2.0, (len(views),) + tuple(shape[1:])).astype(np.float32)
    raw = torch.zeros(shape, dtype=torch.float32, device=dev)
    raw[list(views)] = torch.from_numpy(sample).to(dev)
    return raw, sample


def _central(vol, n):
    return vol[n // 2 - SLAB // 2: n // 2 + SLAB // 2]


def test_full_size_forward_central_rays_and_linearity(dev):
    import ctreconstruction_amd as ctr
    geom = ctr.ScanGeometry(DD_FULL)
    proj = ctr.dd_forward(torch.ones(geom.vol_shape, device=dev), geom)
    # the four cells around the central ray at 0, 90, 180 and 270 degrees cross the whole
    # 128 mm cube (their 0.2 mm offset lengthens the path by under 1e-7)
    c = proj[[0, 180, 360, 540]][:, 383:385, 511:513].cpu().numpy()
    np.testing.assert_allclose(c, 128.0, rtol=2e-6)
    a = torch.rand(geom.vol_shape, device=dev, generator=torch.Generator(device=dev).manual_seed(3))
    lin = ctr.dd_forward(2.0 * a + 1.0, geom)
    torch.testing.assert_close(lin, 2.0 * ctr.dd_forward(a, geom) + proj, rtol=2e-5,
                               atol=2e-5 * float(proj.max()))


def test_full_size_forward_matches_c_oracle_on_eight_views(dev):
    """The benchmark phantom's forward projection at 512^3 / 720 views / 1024 x 768, every cell
    of eight views spread over the scan (view 90 at exactly 45 degrees, where the slab branch
    ties), against the oracle's C restatement of its slab loop (FP64, every host core; the
    oracle's 1e-6 weight zeroing)."""
    import ctreconstruction_amd as ctr
    from ctreconstruction_amd import phantom
    from oracle import dd_c as C

    if not os.path.exists(C.LIB_PATH):
        pytest.fail("oracle/lib/liboracle.so is not built (oracle.dd_c.build())")
    geom = ctr.ScanGeometry(DD_FULL)
    g = O.Geometry(DD_FULL)
    ell = phantom.random_ellipsoids(40, phantom.half_extent(geom), 3)
    vol = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    views = [0, 90, 181, 270, 359, 450, 541, 631]
    got = ctr.dd_forward(vol, geom)[views].cpu().numpy()
    ref = C.forward_project(g, vol.cpu().numpy(), views=views)
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"DD forward 512^3, eight whole views: max rel err {err:.2e}")
    assert err <= TOL


def test_full_size_dd_fdk_matches_oracle_on_two_views(dev):
    import ctreconstruction_amd as ctr
    geom = ctr.ScanGeometry(DD_FULL)
    g = O.Geometry(DD_FULL)
    views = (0, 250)
    raw, sample = _two_view_projections((g.nview, g.nv, g.nu), views, dev, seed=5)
    q = ctr.fdk_filter(raw, geom)
    q_ref = O.filter_projections(g, sample.astype(np.float64))
    err_f = float(np.abs(q[list(views)].cpu().numpy() - q_ref).max() / np.abs(q_ref).max())
    assert float(q[1].abs().max()) == 0.0
    vol = ctr.dd_backproject(q, geom, mode="ref")
    p_slab = dict(DD_FULL, NumberOfImage=[g.nx, g.ny, SLAB])
    gs = O.Geometry(p_slab)
    ref = np.zeros((SLAB, g.ny, g.nx))
    for n, i in enumerate(views):
        ref += O.backproject_view(gs, q_ref[n] * g.dtheta, i, "ref")
    got = _central(vol, g.nz).cpu().numpy()
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"DD FDK 512^3 / 720 views, two-view sample: filter {err_f:.2e}, backprojection {err:.2e}")
    assert err_f <= TOL and err <= TOL


def test_full_size_dd_fdk_matches_c_oracle_on_whole_volume(dev):
    """Eight views spread over the scan of a 40-ellipsoid phantom, filtered and backprojected
    into the whole 512^3 volume on both backprojection grids, against the oracle's C
    restatement (FP64, every host core) fed the same raw projections and the same filtered
    rows respectively."""
    import ctreconstruction_amd as ctr
    from ctreconstruction_amd import phantom
    from oracle import dd_c as C

    if not os.path.exists(C.LIB_PATH):
        pytest.fail("oracle/lib/liboracle.so is not built (oracle.dd_c.build())")
    geom = ctr.ScanGeometry(DD_FULL)
    g = O.Geometry(DD_FULL)
    views = [0, 97, 181, 272, 359, 446, 533, 631]
    sub = geom.with_views(views)
    ell = phantom.random_ellipsoids(40, phantom.half_extent(geom), 3)
    vol = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    raw = ctr.dd_forward(vol, sub)
    del vol
    q = ctr.fdk_filter(raw, geom)
    q_np = q.cpu().numpy()
    q_ref = C.filter_projections(g, raw.cpu().numpy())
    err_f = float(np.abs(q_np - q_ref).max() / np.abs(q_ref).max())
    errs = {}
    for mode in ("ref", "edge"):
        got = ctr.dd_backproject(q, sub, mode=mode).cpu().numpy()
        ref = C.backproject(g, q_np, mode, views=views)
        errs[mode] = float(np.abs(got - ref).max() / np.abs(ref).max())
        del got, ref
    print(f"DD FDK 512^3, eight views into the whole volume: filter {err_f:.2e}, "
          f"backprojection {errs['ref']:.2e} (reference grid), {errs['edge']:.2e} (edge grid)")
    assert err_f <= TOL and max(errs.values()) <= TOL


def test_full_size_complete_fdk_matches_c_oracle(dev):
    """The benchmark's whole reconstruction against the oracle: the 40-ellipsoid phantom's 720
    projections filtered and backprojected into every voxel of the 512^3 volume by the GPU
    (ctr.fdk) and by the oracle's batched filter and C restatement of its backprojection (FP64,
    every host core; about a minute on 16 cores), from the same raw projections."""
    import ctreconstruction_amd as ctr
    from ctreconstruction_amd import phantom
    from oracle import dd_c as C

    if not os.path.exists(C.LIB_PATH):
        pytest.fail("oracle/lib/liboracle.so is not built (oracle.dd_c.build())")
    geom = ctr.ScanGeometry(DD_FULL)
    g = O.Geometry(DD_FULL)
    ell = phantom.random_ellipsoids(40, phantom.half_extent(geom), 3)
    vol = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    proj = ctr.dd_forward(vol, geom)
    del vol
    rec = ctr.fdk(proj, geom).cpu().numpy()
    raw = proj.cpu().numpy()
    del proj
    q_ref = C.filter_projections(g, raw)
    del raw
    ref = C.backproject(g, q_ref, "ref")
    err = float(np.abs(rec - ref).max() / np.abs(ref).max())
    print(f"DD FDK 512^3, all 720 views into the whole volume: max rel err {err:.2e}")
    assert err <= TOL


def test_full_size_pixel_fdk_matches_oracle_on_two_views(dev):
    from ctreconstruction_amd.pixel_fdk import PixelScan, pixel_backproject, pixel_filter
    scan = PixelScan(PIXEL_FULL)
    og = PO.PixelGeometry(PIXEL_FULL)
    views = (0, 170)
    raw, sample = _two_view_projections(scan.proj_shape, views, dev, seed=6)
    q = pixel_filter(raw, scan)
    q_ref = PO.filter_projections(og, sample.astype(np.float64))
    err_f = float(np.abs(q[list(views)].cpu().numpy() - q_ref).max() / np.abs(q_ref).max())
    vol = pixel_backproject(q, scan)
    og.z = _central(og.z, og.nz)
    og.nz = SLAB
    ref = PO.backproject(og, {i: q_ref[n] for n, i in enumerate(views)}, views=views)
    sub = copy.copy(og)
    sub.angles = og.angles[list(views)]
    sub.ns = len(views)
    keep = PO.edge_margin(sub) > 2e-3
    assert keep.mean() > 0.99
    got = _central(vol, scan.nz).cpu().numpy()
    err = float(np.abs(got - ref)[keep].max() / np.abs(ref).max())
    print(f"pixel FDK 512^3 / 680 views, two-view sample: filter {err_f:.2e}, backprojection {err:.2e}")
    assert err_f <= TOL and err <= TOL


# The 256^3 cone-beam FDK workload: 0.5 mm voxels, 360 views, 384 x 384 cells of 1.0 mm, SID 1000,
# SDD 1500, Shepp-Logan (tau = 0.667 mm, so the reference's frequency mask (Q4) is active), a
# phantom of 20 seeded random ellipsoids (seed 2) scaled by mu_water.
DD_C2 = make_params(NumberOfImage=[256, 256, 256], ImagePixelSpacing=[0.5, 0.5, 0.5],
                    NumberOfDetectorPixels=[384, 384], DetectorPixelSize=[1.0, 1.0],
                    NumberOfViews=360, FilterType="shepp-logan")


def _row_band(p, k0, m):
    """The scan of ``p`` seen by detector rows k0 .. k0+m-1 only: a detector of m rows whose v
    offset puts every cell centre and edge where the full detector has it (flat lattice,
    v_k = ((nv-1)/2 - k) dv + off_v)."""
    nu, nv = p["NumberOfDetectorPixels"]
    dv = p["DetectorPixelSize"][1]
    off = list(p.get("DetectorOffset", [0, 0]))
    return dict(p, NumberOfDetectorPixels=[nu, m], DetectorOffset=[off[0], off[1] + ((nv - m) / 2 - k0) * dv])


def test_c2_workload_matches_oracle(dev):
    import ctreconstruction_amd as ctr
    from ctreconstruction_amd import phantom
    geom = ctr.ScanGeometry(DD_C2)
    g = O.Geometry(DD_C2)
    assert np.pi / (2 * geom.tau) < np.pi  # Q4 mask active
    ell = phantom.random_ellipsoids(20, phantom.half_extent(geom), 2)
    vol_t = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    vol = vol_t.cpu().numpy().astype(np.float64)
    proj = ctr.dd_forward(vol_t, geom)
    proj_h = proj.cpu().numpy()
    # forward projection against the oracle on every 30th view, each on one band of 24 rows
    # (the bands cycle through the upper, central and lower part of the volume's shadow)
    e_fp = 0.0
    peak = float(np.abs(proj_h).max())
    for n, i in enumerate(range(0, g.nview, 30)):
        k0 = (100, 180, 260)[n % 3]
        gb = O.Geometry(_row_band(DD_C2, k0, 24))
        ref = O.forward_project_view(gb, vol, i)
        e_fp = max(e_fp, float(np.abs(proj_h[i, k0:k0 + 24] - ref).max()) / peak)
    # the full filter on three views
    q = ctr.fdk_filter(proj, geom)
    views = [0, 123, 271]
    q_ref = O.filter_projections(g, proj_h[views].astype(np.float64))
    e_f = float(np.abs(q[views].cpu().numpy() - q_ref).max() / np.abs(q_ref).max())
    # the backprojection of two filtered views into the central 8 slices
    bp_views = (37, 200)
    q2 = torch.zeros_like(q)
    q2[list(bp_views)] = q[list(bp_views)]
    rec2 = ctr.dd_backproject(q2, geom, mode="ref")
    gs = O.Geometry(dict(DD_C2, NumberOfImage=[g.nx, g.ny, SLAB]))
    ref_bp = np.zeros((SLAB, g.ny, g.nx))
    qh = q.cpu().numpy()
    for i in bp_views:
        ref_bp += O.backproject_view(gs, qh[i].astype(np.float64) * g.dtheta, i, "ref")
    e_bp = float(np.abs(_central(rec2, g.nz).cpu().numpy() - ref_bp).max() / np.abs(ref_bp).max())
    # the whole FDK: inside the phantom's support the reconstruction follows the phantom
    rec = ctr.dd_backproject(q, geom, mode="ref")
    mid = slice(g.nz // 2 - 8, g.nz // 2 + 8)
    rel_rmse = float(torch.sqrt(torch.mean((rec[mid] - vol_t[mid]) ** 2)) / torch.sqrt(torch.mean(vol_t[mid] ** 2)))
    print(f"256^3 / 360 views / 384^2: forward {e_fp:.2e}, filter {e_f:.2e}, backprojection {e_bp:.2e}, "
          f"FDK rel. RMSE vs phantom (central 16 slices) {rel_rmse:.3f}")
    assert e_fp <= TOL and e_f <= TOL and e_bp <= TOL
    assert rel_rmse < 0.5


def test_c2_complete_fdk_matches_c_oracle(dev):
    """The forward projection over every cell of every 30th view, and the whole C2
    reconstruction against the oracle: the phantom's 360 projections filtered and
    backprojected into every voxel of the 256^3 volume by the GPU (ctr.fdk) and by the oracle's
    batched filter and C restatement of its backprojection (FP64, every host core), from the same
    raw projections."""
    import ctreconstruction_amd as ctr
    from ctreconstruction_amd import phantom
    from oracle import dd_c as C

    if not os.path.exists(C.LIB_PATH):
        pytest.fail("oracle/lib/liboracle.so is not built (oracle.dd_c.build())")
    geom = ctr.ScanGeometry(DD_C2)
    g = O.Geometry(DD_C2)
    ell = phantom.random_ellipsoids(20, phantom.half_extent(geom), 2)
    vol_t = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    proj = ctr.dd_forward(vol_t, geom)
    # the forward projection over every cell of every 30th view
    views = list(range(0, g.nview, 30))
    fp_ref = C.forward_project(g, vol_t.cpu().numpy(), views=views)
    e_fp = float(np.abs(proj[views].cpu().numpy() - fp_ref).max() / np.abs(fp_ref).max())
    rec = ctr.fdk(proj, geom).cpu().numpy()
    q_ref = C.filter_projections(g, proj.cpu().numpy())
    ref = C.backproject(g, q_ref, "ref")
    err = float(np.abs(rec - ref).max() / np.abs(ref).max())
    print(f"C2 forward, every cell of 12 views: max rel err {e_fp:.2e}; "
          f"FDK, all 360 views into the whole 256^3 volume: max rel err {err:.2e}")
    assert e_fp <= TOL and err <= TOL


def test_c2_central_rays_and_linearity(dev):
    import ctreconstruction_amd as ctr
    geom = ctr.ScanGeometry(DD_C2)
    proj = ctr.dd_forward(torch.ones(geom.vol_shape, device=dev), geom)
    # the four cells around the central ray of the even-sized detector cross the 128 mm cube
    # (their 0.5 mm offset from the axis lengthens the path by under 2e-7)
    c = proj[[0, 90, 180, 270]][:, 191:193, 191:193].cpu().numpy()
    np.testing.assert_allclose(c, 128.0, rtol=2e-6)
    a = torch.rand(geom.vol_shape, device=dev, generator=torch.Generator(device=dev).manual_seed(4))
    lin = ctr.dd_forward(2.0 * a + 1.0, geom)
    torch.testing.assert_close(lin, 2.0 * ctr.dd_forward(a, geom) + proj, rtol=2e-5,
                               atol=2e-5 * float(proj.max()))


def test_full_size_pixel_fdk_matches_c_oracle_on_whole_volume(dev):
    """Eight views spread over the 680-view scan of a 40-ellipsoid phantom, filtered and
    backprojected into the whole 512^3 volume by the GPU and by the pixel oracle's C restatement
    (FP64, every host core), compared on the voxels whose sample stays 2e-3 cells away from a
    bounds-test flip in every view."""
    import ctreconstruction_amd as ctr
    from ctreconstruction_amd import paramsfile, phantom
    from ctreconstruction_amd.pixel_fdk import PixelScan, pixel_backproject, pixel_filter
    from oracle import dd_c as C
    from oracle import pixel_c as PC

    if not os.path.exists(C.LIB_PATH):
        pytest.fail("oracle/lib/liboracle.so is not built (oracle.dd_c.build())")
    scan = PixelScan(PIXEL_FULL)
    og = PO.PixelGeometry(PIXEL_FULL)
    views = [0, 83, 171, 254, 340, 423, 509, 597]
    geom = ctr.ScanGeometry(paramsfile.to_recon_params(PIXEL_FULL))
    ell = phantom.random_ellipsoids(40, phantom.half_extent(geom), 4)
    vol = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    raw = ctr.dd_forward(vol, geom.with_views(views))
    del vol
    q = pixel_filter(raw, scan)
    q_np = q.cpu().numpy()
    q_ref = PC.filter_projections(og, raw.cpu().numpy())
    err_f = float(np.abs(q_np - q_ref).max() / np.abs(q_ref).max())
    sub = PixelScan({**PIXEL_FULL, "NumberOfViews": len(views)})
    ang = scan.device_angles(dev)[views].contiguous()
    got = pixel_backproject(q, sub, scale=scan.dtheta, angles=ang).cpu().numpy()
    ref, margin = PC.backproject(og, q_np, views, margin=True)
    keep = margin > 2e-3
    assert keep.mean() > 0.99
    err = float(np.abs(got - ref)[keep].max() / np.abs(ref).max())
    print(f"pixel FDK 512^3, eight views into the whole volume: filter {err_f:.2e}, "
          f"backprojection {err:.2e} ({keep.mean():.4f} of the voxels)")
    assert err_f <= TOL and err <= TOL


def test_full_size_pixel_fdk_all_views_match_c_oracle_on_central_slab(dev):
    """Every one of the 680 views of a 40-ellipsoid phantom scan, weighted, filtered and
    backprojected by the GPU, against the pixel oracle's batched filter and C backprojection
    (FP64, every host core) over the central 64 slices of the 512^3 volume, on the voxels whose
    sample stays 2e-3 cells away from a bounds-test flip in every view."""
    import ctreconstruction_amd as ctr
    from ctreconstruction_amd import paramsfile, phantom
    from ctreconstruction_amd.pixel_fdk import PixelScan, pixel_fdk
    from oracle import dd_c as C
    from oracle import pixel_c as PC

    if not os.path.exists(C.LIB_PATH):
        pytest.fail("oracle/lib/liboracle.so is not built (oracle.dd_c.build())")
    scan = PixelScan(PIXEL_FULL)
    og = PO.PixelGeometry(PIXEL_FULL)
    geom = ctr.ScanGeometry(paramsfile.to_recon_params(PIXEL_FULL))
    ell = phantom.random_ellipsoids(40, phantom.half_extent(geom), 5)
    vol = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    raw = ctr.dd_forward(vol, geom)
    del vol
    slab = 64
    mid = slice(og.nz // 2 - slab // 2, og.nz // 2 + slab // 2)
    got = pixel_fdk(raw, scan)[mid].cpu().numpy()
    q_ref = PC.filter_projections(og, raw.cpu().numpy())
    del raw
    og.z = og.z[mid]
    og.nz = slab
    ref, margin = PC.backproject(og, q_ref, range(og.ns), margin=True)
    keep = margin > 2e-3
    assert keep.mean() > 0.99
    err = float(np.abs(got - ref)[keep].max() / np.abs(ref).max())
    print(f"pixel FDK 512^3, all 680 views into the central {slab} slices: max rel err {err:.2e} "
          f"({keep.mean():.4f} of the voxels)")
    assert err <= TOL
