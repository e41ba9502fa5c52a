"""GPU parity: the HIP kernels against the FP64 oracle on the same seeded inputs.

Tolerance (from the task's parity bar): max |gpu - oracle| / max |oracle| <= 1e-4 for every
result; the measured figure is printed per case.  Sizes are small enough for the oracle to
finish in seconds.
"""
import numpy as np
import pytest
import torch

import ctreconstruction_amd as ctr
from ctreconstruction_amd import _native as N
from oracle import dd_oracle as O
from tests.conftest import make_params

pytestmark = pytest.mark.gpu

TOL = 1e-4

CASES = {
    "flat": dict(NumberOfImage=[24, 24, 20], NumberOfDetectorPixels=[40, 32], NumberOfViews=12),
    "flat_45deg_ties": dict(NumberOfImage=[20, 20, 16], NumberOfDetectorPixels=[36, 28], NumberOfViews=8),
    "flat_offsets": dict(NumberOfImage=[22, 18, 14], NumberOfDetectorPixels=[37, 29], NumberOfViews=10,
                         StartAngle=0.3, EndAngle=0.3 + 2 * np.pi, ReconCenter=[0.7, -0.4, 0.3],
                         DetectorOffset=[1.3, -0.6]),
    "anisotropic": dict(NumberOfImage=[26, 18, 12], ImagePixelSpacing=[0.4, 0.6, 0.8],
                        NumberOfDetectorPixels=[45, 23], DetectorPixelSize=[0.9, 1.3], NumberOfViews=9,
                        StartAngle=-0.2, EndAngle=np.pi),
    "helical": dict(NumberOfImage=[20, 20, 24], NumberOfDetectorPixels=[36, 20], NumberOfViews=16,
                    Pitch=0.5, EndAngle=4 * np.pi),
    # a fine angular step: the forward projector walks each pair of adjacent views together
    # (one staged segment for both), including pairs that straddle a change of slab orientation
    "fine_steps": dict(NumberOfImage=[24, 24, 20], NumberOfDetectorPixels=[40, 32], NumberOfViews=90,
                       StartAngle=0.01, EndAngle=0.01 + 2 * np.pi),
    "curved": dict(NumberOfImage=[24, 24, 16], NumberOfDetectorPixels=[48, 30], NumberOfViews=12,
                   DetectorShape="Curved"),
    # 2 mm voxels seen by 0.2 mm pixels: more u taps than are held in registers, detector windows
    # larger than the LDS staging buffer, and voxels projecting outside the detector
    "wide_footprint": dict(NumberOfImage=[20, 20, 10], ImagePixelSpacing=[2.0, 2.0, 2.0],
                           NumberOfDetectorPixels=[300, 150], DetectorPixelSize=[0.2, 0.2], NumberOfViews=6,
                           StartAngle=0.1, EndAngle=0.1 + 2 * np.pi),
    "curved_fan_2d": dict(NumberOfImage=[48, 48, 1], ImagePixelSpacing=[0.5, 0.5, 0.5],
                          NumberOfDetectorPixels=[96, 1], DetectorPixelSize=[0.5, 0.75], NumberOfViews=24,
                          DetectorShape="Curved"),
    # a column wavefront's 64 row edges span 83..166 z cells at most slabs: the forward projector
    # stages its z segment with two loads per slab
    "tall_rows": dict(NumberOfImage=[20, 20, 200], NumberOfDetectorPixels=[36, 128],
                      DetectorPixelSize=[1.0, 1.05], NumberOfViews=8),
    # ... more than 166 cells: the forward projector's explicit per-line path
    "coarse_rows": dict(NumberOfImage=[12, 12, 256], NumberOfDetectorPixels=[24, 72],
                        DetectorPixelSize=[1.2, 2.4], NumberOfViews=6),
    # a volume far off the rotation axis (shifted by about half its width in x, y and z)
    "off_center": dict(NumberOfImage=[24, 20, 18], ReconCenter=[6.0, -5.0, 4.0], NumberOfDetectorPixels=[48, 40],
                       NumberOfViews=12, StartAngle=0.2, EndAngle=0.2 + 2 * np.pi),
    # the source 60 mm from the isocentre: a voxel column's footprint height changes by ~2.4 %
    # per voxel along z, too fast for the backprojector's series normaliser, so its
    # sub-chunks take exact reciprocals in the completing pass
    "near_source": dict(SourceInit=[0, 60.0, 0], DetectorInit=[0, -40.0, 0], NumberOfImage=[16, 16, 32],
                        ImagePixelSpacing=[1.0, 1.0, 1.0], NumberOfDetectorPixels=[20, 40],
                        DetectorPixelSize=[2.0, 2.0], NumberOfViews=12, StartAngle=0.1,
                        EndAngle=0.1 + 2 * np.pi),
    # rotation in the negative sense: decreasing angles, dtheta < 0 (the reference's linspace
    # allows it, and Q = proj * dtheta changes sign with it)
    "reverse": dict(NumberOfImage=[24, 24, 20], NumberOfDetectorPixels=[40, 32], NumberOfViews=12,
                    StartAngle=0.1 + 2 * np.pi, EndAngle=0.1),
    "helical_reverse": dict(NumberOfImage=[20, 20, 24], NumberOfDetectorPixels=[36, 20], NumberOfViews=16,
                            Pitch=0.5, StartAngle=4 * np.pi, EndAngle=0.0),
    "short_scan": dict(NumberOfImage=[24, 22, 18], NumberOfDetectorPixels=[40, 32], NumberOfViews=30,
                       StartAngle=0.05, EndAngle=0.05 + np.pi + 0.6),
    # source and detector centre at different heights: the central ray is tilted
    "source_z": dict(SourceInit=[0, 1000.0, 6.0], DetectorInit=[0, -500.0, -3.0], NumberOfImage=[24, 24, 20],
                     NumberOfDetectorPixels=[40, 36], NumberOfViews=12, StartAngle=0.1, EndAngle=0.1 + 2 * np.pi),
    # SAD measured from a rotation origin off the source's axis (900 mm instead of 1000)
    "origin": dict(RotationOrigin=[0, 100.0, 0], NumberOfImage=[24, 24, 20], NumberOfDetectorPixels=[40, 32],
                   NumberOfViews=12, StartAngle=0.1, EndAngle=0.1 + 2 * np.pi),
    # one view: dtheta is EndAngle - StartAngle
    "single_view": dict(NumberOfImage=[24, 24, 20], NumberOfDetectorPixels=[40, 32], NumberOfViews=1,
                        StartAngle=0.7, EndAngle=0.7 + 2 * np.pi),
    "thin_z": dict(NumberOfImage=[24, 24, 2], NumberOfDetectorPixels=[40, 32], NumberOfViews=12,
                   StartAngle=0.1, EndAngle=0.1 + 2 * np.pi),
    # the curved detector with the other features (its u offset is an angle, in radians)
    "curved_helical": dict(NumberOfImage=[20, 20, 24], NumberOfDetectorPixels=[48, 24], NumberOfViews=16,
                           Pitch=0.5, EndAngle=4 * np.pi, DetectorShape="Curved"),
    "curved_offsets": dict(NumberOfImage=[22, 18, 14], NumberOfDetectorPixels=[47, 29], NumberOfViews=10,
                           StartAngle=0.3, EndAngle=0.3 + 2 * np.pi, ReconCenter=[0.7, -0.4, 0.3],
                           DetectorOffset=[0.02, -0.6], DetectorShape="Curved"),
    "curved_reverse": dict(NumberOfImage=[24, 24, 16], NumberOfDetectorPixels=[48, 30], NumberOfViews=12,
                           StartAngle=0.1 + 2 * np.pi, EndAngle=0.1, DetectorShape="Curved"),
    "curved_near_source": dict(SourceInit=[0, 60.0, 0], DetectorInit=[0, -40.0, 0], NumberOfImage=[16, 16, 32],
                               ImagePixelSpacing=[1.0, 1.0, 1.0], NumberOfDetectorPixels=[40, 40],
                               DetectorPixelSize=[2.0, 2.0], NumberOfViews=12, StartAngle=0.1,
                               EndAngle=0.1 + 2 * np.pi, DetectorShape="Curved"),
    "aniso_helical": dict(NumberOfImage=[26, 18, 20], ImagePixelSpacing=[0.4, 0.6, 0.5],
                          NumberOfDetectorPixels=[45, 23], DetectorPixelSize=[0.9, 1.3], NumberOfViews=18,
                          Pitch=0.7, StartAngle=-0.2, EndAngle=4 * np.pi - 0.2),
    "flat_fan_2d": dict(NumberOfImage=[48, 48, 1], ImagePixelSpacing=[0.5, 0.5, 0.5],
                        NumberOfDetectorPixels=[96, 1], DetectorPixelSize=[0.5, 0.75], NumberOfViews=24),
}


def _edge_span(p, rows=64):
    """z cells spanned at the isocentre by `rows` consecutive row edges of the geometry, and
    whether the volume's z extent covers every detector row (no clamping shortens the span)."""
    mag = (p["SourceInit"][1] - p["DetectorInit"][1]) / p["SourceInit"][1]
    dv, dz = p["DetectorPixelSize"][1], p["ImagePixelSpacing"][2]
    covered = p["NumberOfImage"][2] * dz * mag > p["NumberOfDetectorPixels"][1] * dv
    return (rows - 1) * dv / (mag * dz), covered


def test_segment_cases_reach_their_paths():
    span, covered = _edge_span(make_params(**CASES["tall_rows"]))
    assert covered and 82 < span < 166
    span, covered = _edge_span(make_params(**CASES["coarse_rows"]))
    assert covered and span > 166


def _rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def _vol(g, seed):
    return np.random.default_rng(seed).random((g.nz, g.ny, g.nx)).astype(np.float32)


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    N.lib()
    return torch.device("cuda:0")


@pytest.mark.parametrize("name", sorted(CASES))
def test_forward_matches_oracle(dev, name):
    p = make_params(**CASES[name])
    g = O.Geometry(p)
    vol = _vol(g, 11)
    out = ctr.dd_forward(torch.from_numpy(vol).to(dev), ctr.ScanGeometry(p)).cpu().numpy()
    ref = O.forward_project(g, vol.astype(np.float64))
    err = _rel(out, ref)
    print(f"forward[{name}] max rel err {err:.2e}")
    assert err <= TOL


def test_forward_into_unaligned_output_view(dev):
    """An output view 4 bytes off a 16-byte boundary (nu a multiple of 4): the row stores fall
    back from 16-byte to per-column writes, with the same values, and nothing outside the view
    is written."""
    p = make_params(**CASES["flat"])
    geom = ctr.ScanGeometry(p)
    vol = torch.from_numpy(_vol(O.Geometry(p), 11)).to(dev)
    ref = ctr.dd_forward(vol, geom)
    n = ref.numel()
    buf = torch.full((n + 2,), -7.0, dtype=torch.float32, device=dev)
    view = buf[1:n + 1].view(ref.shape)
    assert view.data_ptr() % 16 == 4
    ctr.dd_forward(vol, geom, out=view)
    torch.testing.assert_close(view, ref, rtol=0, atol=0)
    assert float(buf[0]) == -7.0 and float(buf[-1]) == -7.0


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("mode", ["ref", "edge"])
def test_backproject_matches_oracle(dev, name, mode):
    p = make_params(**CASES[name])
    g = O.Geometry(p)
    q = np.random.default_rng(12).standard_normal((g.nview, g.nv, g.nu)).astype(np.float32)
    out = ctr.dd_backproject(torch.from_numpy(q).to(dev), ctr.ScanGeometry(p), mode=mode).cpu().numpy()
    ref = O.backproject(g, q, mode)
    err = _rel(out, ref)
    print(f"backproject[{name},{mode}] max rel err {err:.2e}")
    assert err <= TOL


def test_forward_z_slab_views_match_oracle(dev):
    """Views that select the z-slab branch (steep rays; the oracle's z branch is pinned by brute
    force and a uniform-volume closed form) next to x- and y-slab views of the same scan."""
    from tests.test_oracle_projectors import ZSLAB
    p = make_params(**ZSLAB)
    g = O.Geometry(p)
    geom = ctr.ScanGeometry(p)
    assert list(geom.branch) == [1, 2, 0, 2, 1, 2, 0, 2]
    vol = _vol(g, 17)
    out = ctr.dd_forward(torch.from_numpy(vol).to(dev), geom).cpu().numpy()
    ref = O.forward_project(g, vol.astype(np.float64))
    zv = geom.branch == N.CTR_BRANCH_Z
    e_z, e_xy = _rel(out[zv], ref[zv]), _rel(out[~zv], ref[~zv])
    print(f"forward z-slab views {e_z:.2e}, x/y-slab views {e_xy:.2e}")
    assert np.abs(ref[zv]).max() > 0
    assert e_z <= TOL and e_xy <= TOL
    # the matched adjoint covers x- and y-slab views only, and says so
    with pytest.raises(N.CtrError):
        ctr.dd_backproject(torch.from_numpy(out).to(dev), geom, mode="adjoint")


def test_backproject_numba_weight_zeroing(dev):
    # the numba kernels zero overlap weights below 1e-5 (python/GPUFuncs_numba.py:28-29, 45-46)
    p = make_params(**CASES["flat"])
    g = O.Geometry(p)
    q = np.random.default_rng(13).random((g.nview, g.nv, g.nu)).astype(np.float32)
    out = ctr.dd_backproject(torch.from_numpy(q).to(dev), ctr.ScanGeometry(p), zero_tol=1e-5).cpu().numpy()
    assert _rel(out, O.backproject(g, q, "ref", zero_tol=1e-5)) <= TOL


@pytest.mark.parametrize("ftype", O.FILTER_TYPES)
@pytest.mark.parametrize("du", [1.0, 0.4])
def test_filter_spectrum_matches_oracle(dev, ftype, du):
    p = make_params(NumberOfDetectorPixels=[384, 8], DetectorPixelSize=[du, du], FilterType=ftype, cutoff=0.9)
    geom = ctr.ScanGeometry(p)
    spec = ctr.filter_spectrum(geom, device=dev).cpu().numpy()
    ref = np.fft.ifftshift(O.ramp_spectrum(geom.zero_padded_length + 1, geom.tau, ftype, 0.9))
    np.testing.assert_allclose(spec, ref, rtol=0, atol=1e-12 * ref.max())


@pytest.mark.parametrize("ftype", O.FILTER_TYPES)
@pytest.mark.parametrize("nu", [37, 384, 1024])
def test_fdk_filter_matches_oracle(dev, ftype, nu):
    p = make_params(NumberOfDetectorPixels=[nu, 7], NumberOfViews=3, FilterType=ftype, cutoff=0.8,
                    DetectorOffset=[0.25, -0.5])
    g = O.Geometry(p)
    proj = np.random.default_rng(14).random((3, 7, nu)).astype(np.float32)
    out = ctr.fdk_filter(torch.from_numpy(proj).to(dev), ctr.ScanGeometry(p)).cpu().numpy()
    ref = O.filter_projections(g, proj)
    err = _rel(out, ref)
    print(f"filter[{ftype},{nu}] max rel err {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("nu", [5, 9, 17, 33, 2049, 4097])
def test_fdk_filter_every_fft_shape(dev, nu):
    """FFT lengths 8 (one radix-8 stage), 16 (one radix-16 stage), 32 and 64 (radix 16 then 2
    or 4), 4096 and 8192 (three radix-16 stages, then none or radix 2, on 256-thread groups)."""
    p = make_params(NumberOfDetectorPixels=[nu, 3], NumberOfViews=2, FilterType="shepp-logan",
                    DetectorOffset=[0.25, -0.5])
    g = O.Geometry(p)
    proj = np.random.default_rng(16).random((2, 3, nu)).astype(np.float32)
    out = ctr.fdk_filter(torch.from_numpy(proj).to(dev), ctr.ScanGeometry(p)).cpu().numpy()
    err = _rel(out, O.filter_projections(g, proj))
    print(f"filter[nu={nu}] max rel err {err:.2e}")
    assert err <= TOL


def test_fdk_filter_in_place_and_odd_row_count(dev):
    p = make_params(NumberOfDetectorPixels=[64, 5], NumberOfViews=3)
    g = O.Geometry(p)
    proj = np.random.default_rng(15).random((3, 5, 64)).astype(np.float32)
    t = torch.from_numpy(proj).to(dev)
    ctr.fdk_filter(t, ctr.ScanGeometry(p), out=t)
    assert _rel(t.cpu().numpy(), O.filter_projections(g, proj)) <= TOL


@pytest.mark.parametrize("name", ["flat", "curved", "reverse", "short_scan", "source_z", "single_view"])
def test_fdk_end_to_end_matches_oracle(dev, name):
    p = make_params(FilterType="shepp-logan", **CASES[name])
    g = O.Geometry(p)
    vol = _vol(g, 16)
    geom = ctr.ScanGeometry(p)
    proj = ctr.dd_forward(torch.from_numpy(vol).to(dev), geom)
    rec = ctr.fdk(proj, geom).cpu().numpy()
    p_ref = O.forward_project(g, vol.astype(np.float64))
    ref = O.fdk(g, p_ref.astype(np.float32))
    err = _rel(rec, ref)
    print(f"fdk[{name}] max rel err {err:.2e}")
    assert err <= TOL


def test_results_are_deterministic(dev):
    p = make_params(**CASES["flat_offsets"])
    geom = ctr.ScanGeometry(p)
    vol = torch.from_numpy(_vol(O.Geometry(p), 17)).to(dev)
    a = ctr.dd_forward(vol, geom)
    b = ctr.dd_forward(vol, geom)
    assert torch.equal(a, b)
    ra = ctr.fdk(a, geom)
    rb = ctr.fdk(b, geom)
    assert torch.equal(ra, rb)


def test_accumulate_adds_to_existing_volume(dev):
    p = make_params(**CASES["flat"])
    geom = ctr.ScanGeometry(p)
    q = torch.rand(geom.proj_shape, device=dev)
    once = ctr.dd_backproject(q, geom)
    acc = once.clone()
    ctr.dd_backproject(q, geom, out=acc, accumulate=True)
    torch.testing.assert_close(acc, 2 * once, rtol=1e-6, atol=0)


def test_reconstruction_class_drop_in(dev, tmp_path):
    # FDK_512.py flow: LoadRecon -> forward -> SaveProj -> Filtering -> backward -> SaveRecon
    p = make_params(FilterType="hann", **CASES["flat"])
    g = O.Geometry(p)
    vol = _vol(g, 18)
    vol.tofile(tmp_path / "phantom.dat")
    R = ctr.Reconstruction(p)
    R.LoadRecon(str(tmp_path / "phantom.dat"))
    R.forward()
    assert isinstance(R.proj, np.ndarray) and R.proj.shape == (g.nview, g.nv, g.nu)
    R.SaveProj(str(tmp_path / "proj.dat"))
    p_ref = O.forward_project(g, vol.astype(np.float64))
    assert _rel(np.fromfile(tmp_path / "proj.dat", np.float32).reshape(R.proj.shape), p_ref) <= TOL
    R.Filtering()
    R.backward()
    R.SaveRecon(str(tmp_path / "recon.dat"))
    ref = O.fdk(g, p_ref.astype(np.float32))
    assert _rel(np.fromfile(tmp_path / "recon.dat", np.float32).reshape(g.nz, g.ny, g.nx), ref) <= TOL


def test_reconstruction_static_filter_matches_oracle(dev):
    for ftype in O.FILTER_TYPES:
        H = ctr.Reconstruction.Filter(1025, 0.6667, ftype, 1)
        np.testing.assert_allclose(H, O.ramp_spectrum(1025, 0.6667, ftype, 1.0), rtol=0, atol=1e-12 * H.max())


def test_uniform_volume_central_ray_at_c2_size(dev):
    # size-independent property at the 256^3 / 384^2 / 360-view configuration:
    # a uniform volume's central ray equals its length (ny * dy), views 0 and 90 degrees.
    p = make_params(NumberOfImage=[256, 256, 256], NumberOfDetectorPixels=[385, 385], NumberOfViews=360,
                    ImagePixelSpacing=[0.5, 0.5, 0.5], DetectorPixelSize=[1.0, 1.0])
    geom = ctr.ScanGeometry(p)
    proj = ctr.dd_forward(torch.ones(geom.vol_shape, device=dev), geom)
    c = proj[:, 192, 192].cpu().numpy()
    np.testing.assert_allclose(c[[0, 90, 180, 270]], 128.0, rtol=2e-6)
    # linearity of the projector at full size
    a = torch.rand(geom.vol_shape, device=dev)
    lin = ctr.dd_forward(2.0 * a + 1.0, geom)
    torch.testing.assert_close(lin, 2.0 * ctr.dd_forward(a, geom) + proj, rtol=2e-5, atol=2e-5 * float(proj.max()))


def _c1_params():
    """The 2D equiangular fan-beam case at full size: 256^2 at 0.5 mm, 360 views over 360
    degrees, 512 bins on an arc with 0.5/1500 rad spacing, nz = nv = 1, SID 1000, SDD 1500."""
    return make_params(NumberOfImage=[256, 256, 1], ImagePixelSpacing=[0.5, 0.5, 0.5],
                       NumberOfDetectorPixels=[512, 1], DetectorPixelSize=[0.5, 0.5], NumberOfViews=360,
                       DetectorShape="Curved", FilterType="ram-lak")


def test_fan_2d_full_size_fbp_matches_oracle(dev):
    """Forward projection of a 10-ellipse phantom (seed 1), ramp filter and DD backprojection
    (reference grid and true cell edges) at the full 2D fan-beam size, against the oracle."""
    from ctreconstruction_amd import phantom
    p = _c1_params()
    g = O.Geometry(p)
    geom = ctr.ScanGeometry(p)
    ell = phantom.random_ellipsoids(10, phantom.half_extent(geom), 1, ndim=2)
    vol_t = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    vol = vol_t.cpu().numpy().astype(np.float64)
    proj = ctr.dd_forward(vol_t, geom)
    views = list(range(0, g.nview, 6))  # the oracle's forward projection on every sixth view
    ref_p = O.forward_project(g, vol, views=views)
    e_fp = _rel(proj.cpu().numpy()[views], ref_p)
    q = ctr.fdk_filter(proj, geom)
    ref_q = O.filter_projections(g, proj.cpu().numpy().astype(np.float64))
    e_f = _rel(q.cpu().numpy(), ref_q)
    errs = {"forward": e_fp, "filter": e_f}
    for mode in ("ref", "edge"):
        rec = ctr.dd_backproject(q, geom, mode=mode).cpu().numpy()
        ref_r = O.backproject(g, q.cpu().numpy().astype(np.float64), mode)
        errs["fbp_" + mode] = _rel(rec, ref_r)
    print("fan 2D full size, max rel err:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) <= TOL


def test_legacy_forward_backward_entry_points(dev, tmp_path):
    """Forward(filename, params).forward() and Backward(filename, params).backward() read raw
    float32 files and give the oracle's projection / filtered edge-grid backprojection on the
    legacy voxel grid."""
    from ctreconstruction_amd.reconstruction import _legacy_params
    p = make_params(NumberOfImage=[20, 18, 12], NumberOfDetectorPixels=[40, 28], NumberOfViews=10,
                    FilterType="hann")
    p["PhantomCenter"] = [1.0, -1.0, 0.5]
    p["Origin"] = [0, 0, 0]
    g = O.Geometry(_legacy_params(p))
    vol = _vol(g, 13)
    f_vol = tmp_path / "vol.dat"
    vol.tofile(f_vol)
    fw = ctr.Forward(str(f_vol), p)
    fw.forward()
    ref = O.forward_project(g, vol.astype(np.float64))
    e_f = _rel(fw.proj, ref)
    f_proj = tmp_path / "proj.dat"
    np.asarray(fw.proj, np.float32).tofile(f_proj)
    bw = ctr.Backward(str(f_proj), p)
    bw.backward()
    q = O.filter_projections(g, np.asarray(fw.proj, np.float64))
    ref_b = O.backproject(g, q, "edge")
    e_b = _rel(bw.recon, ref_b)
    print(f"legacy entry points: forward {e_f:.2e}, backward {e_b:.2e}")
    assert e_f <= TOL and e_b <= TOL


@pytest.mark.parametrize("mode", ["edge", "ref"])
def test_fdk_of_forward_projection_puts_blobs_in_place(dev, mode):
    """fdk(dd_forward(x)) of three off-centre, mutually asymmetric ellipsoids at 64^3 puts each
    one's centre of mass where it is: ties the forward projector's detector lattice on the GPU to
    the backprojector's orientation, which the oracle pins against exact line integrals.  The
    reference grid (Q1) moves the object by -dv/2 * SAD/SDD along z."""
    from tests import analytic
    p = make_params(NumberOfImage=[64, 64, 64], ImagePixelSpacing=[0.5, 0.5, 0.5],
                    NumberOfDetectorPixels=[96, 88], DetectorPixelSize=[0.8, 0.8], NumberOfViews=90,
                    FilterType="ram-lak")
    g = O.Geometry(p)
    ell = analytic.asymmetric_blobs(16.0)
    xs, ys, zs = g.centres()
    vol = analytic.voxelise(ell, xs, ys, zs, sub=2).astype(np.float32)
    geom = ctr.ScanGeometry(p)
    rec = ctr.fdk(ctr.dd_forward(torch.from_numpy(vol).to(dev), geom), geom, mode=mode).cpu().numpy()
    err = analytic.centres_of_mass(rec.astype(np.float64), xs, ys, zs, ell) - np.array([e[0] for e in ell])
    if mode == "ref":
        err[:, 2] += 0.5 * g.dv * g.sad / g.sdd
    print(f"blob centres [{mode}], error in voxels:\n{err / g.dx}")
    assert np.abs(err[:, :2]).max() < 0.25 * g.dx
    assert np.abs(err[:, 2]).max() < (0.25 * g.dz if mode == "edge" else 0.1)


@pytest.mark.parametrize("name", ["flat", "wide_footprint", "helical"])
def test_forward_against_both_reference_weight_zeroings(dev, name):
    """The pycuda kernels zero overlap weights below 1e-6 in both directions
    (python/GPUFuncs_pycuda.py:434-449), the numba kernels below 1e-5 (python/GPUFuncs_numba.py:16-47).
    The z-prefix projector zeroes transverse weights below 1e-6 and never splits z overlaps per
    cell, so its deviation from each variant (and from no zeroing) is measured here."""
    p = make_params(**CASES[name])
    g = O.Geometry(p)
    vol = _vol(g, 21)
    out = ctr.dd_forward(torch.from_numpy(vol).to(dev), ctr.ScanGeometry(p)).cpu().numpy()
    errs = {tol: _rel(out, O.forward_project(g, vol.astype(np.float64), zero_tol=tol))
            for tol in (1e-6, 1e-5, None)}
    print(f"forward[{name}] vs oracle zeroing 1e-6 / 1e-5 / none: "
          + " / ".join(f"{e:.2e}" for e in errs.values()))
    assert max(errs.values()) <= TOL


def test_numba_variant_planes_on_phantom_centre(dev):
    """Reconstruction(params, variant="numba"): the forward projector's planes sit on PhantomCenter
    (python/Reconstruction_numba.py:835-837), the backprojection stays on ReconCenter with the
    numba kernels' 1e-5 weight zeroing (python/GPUFuncs_numba.py:28-29, 45-46)."""
    kw = dict(CASES["flat"], ReconCenter=[0.5, -0.3, 0.2], FilterType="ram-lak")
    p = make_params(PhantomCenter=[-1.0, 0.75, -0.5], **kw)
    g_fp = O.Geometry(make_params(**dict(kw, ReconCenter=[-1.0, 0.75, -0.5])))
    g_bp = O.Geometry(p)
    vol = _vol(g_bp, 22)
    R = ctr.Reconstruction(p, variant="numba")
    R.image = vol
    R.forward()
    ref_p = O.forward_project(g_fp, vol.astype(np.float64))
    e_f = _rel(R.proj, ref_p)
    R.backward()  # unfiltered: the backprojector alone
    ref_b = O.backproject(g_bp, np.asarray(R.proj, np.float64), "ref", zero_tol=1e-5)
    e_b = _rel(R.image, ref_b)
    # the default (pycuda) variant keeps ReconCenter for both
    R2 = ctr.Reconstruction(p)
    R2.image = vol
    R2.forward()
    e_p = _rel(R2.proj, O.forward_project(g_bp, vol.astype(np.float64)))
    print(f"numba variant: forward {e_f:.2e}, backprojection {e_b:.2e}; pycuda variant forward {e_p:.2e}")
    assert e_f <= TOL and e_b <= TOL and e_p <= TOL
    assert _rel(R.proj, O.forward_project(g_bp, vol.astype(np.float64))) > 10 * TOL


# Volumes of 48 slices (a 16 x 16 x 32 voxel tile of two 16-voxel sub-chunks over a tile of one)
# whose (tile, view) pairs fall partly outside the backprojector's common case, so that the
# completing pass has work:
SPLIT_CASES = {
    # a sub-chunk spans 16 * 0.5 * mag / 0.63 = 18.8 .. 19.3 rows (+1 for its partial end rows):
    # some sub-chunks fit the 19-row limit of the row-integral scratch, some do not
    "rows_over_19": dict(NumberOfImage=[32, 32, 48], NumberOfDetectorPixels=[40, 80],
                         DetectorPixelSize=[1.0, 0.63], NumberOfViews=10, StartAngle=0.05,
                         EndAngle=0.05 + 2 * np.pi),
    # a voxel covers 0.5 * mag / 0.3 = 2.5 (axis-aligned) .. 3.5 (diagonal) cells: more than 3
    # u taps at some views; a 16 x 16 x 32 tile's window is 40 .. 57 cells (+2 of margin) by
    # about 17 rows (+2), so some windows fit the 1024-cell LDS buffer and some do not
    "u_taps_and_windows": dict(NumberOfImage=[32, 32, 48], NumberOfDetectorPixels=[120, 32],
                               DetectorPixelSize=[0.3, 1.4], NumberOfViews=10, StartAngle=0.05,
                               EndAngle=0.05 + 2 * np.pi),
}


def test_split_cases_reach_the_completing_pass():
    mag_lo, mag_hi = 1500.0 / (1000.0 + 8 * np.sqrt(2)), 1500.0 / (1000.0 - 8 * np.sqrt(2))
    p = SPLIT_CASES["rows_over_19"]
    rows = 16 * 0.5 / p["DetectorPixelSize"][1]
    assert rows * mag_lo < 19 < rows * mag_hi + 1
    p = SPLIT_CASES["u_taps_and_windows"]
    du = p["DetectorPixelSize"][0]
    assert 0.5 * mag_lo / du < 3 < 0.5 * np.sqrt(2) * mag_hi / du
    rows = 32 * 0.5 * mag_hi / p["DetectorPixelSize"][1] + 2
    assert (8 * mag_lo / du + 2) * rows < 1024 < (8 * np.sqrt(2) * mag_hi / du + 2) * rows


@pytest.mark.parametrize("name", sorted(SPLIT_CASES) + ["flat", "near_source"])
@pytest.mark.parametrize("mode", ["ref", "edge"])
def test_two_pass_backprojection_equals_one_pass(dev, name, mode):
    """The common-case pass plus the completing pass add every (voxel, view) pair exactly once:
    they agree with the single pass over all cases (zero_tol < 0) up to the order of the sums,
    and both agree with the oracle."""
    kw = dict(CASES.get(name, SPLIT_CASES.get(name)))
    if name == "flat":
        kw["NumberOfImage"] = [24, 24, 48]
    p = make_params(**kw)
    g = O.Geometry(p)
    geom = ctr.ScanGeometry(p)
    q = torch.from_numpy(np.random.default_rng(23).random((g.nview, g.nv, g.nu)).astype(np.float32)).to(dev)
    two = ctr.dd_backproject(q, geom, mode=mode).cpu().numpy()
    one = ctr.dd_backproject(q, geom, mode=mode, zero_tol=-1.0).cpu().numpy()
    ref = O.backproject(g, q.cpu().numpy(), mode)
    e_split = float(np.abs(two - one).max() / np.abs(one).max())
    e_ref = _rel(two, ref)
    print(f"backprojection split [{name},{mode}]: two-pass vs one-pass {e_split:.2e}, vs oracle {e_ref:.2e}")
    assert e_split <= 1e-6
    assert e_ref <= TOL
