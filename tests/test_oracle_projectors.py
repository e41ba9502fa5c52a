"""Pins the oracle's DD forward projector and FDK backprojector.

Checks are independent of the oracle's own formulas: closed forms for a uniform
box, analytic ellipsoid line integrals, 90-degree rotation symmetry between the
x- and y-slab branches, brute-force scalar loops, the adjoint identity against
an explicit matrix, and the FDK reconstruction of an analytic sphere.
"""
import numpy as np
import pytest

from oracle import brute as B
from oracle import dd_oracle as O
from tests import analytic
from tests.conftest import make_params


def test_uniform_box_central_ray_equals_path_length():
    g = O.Geometry(make_params())
    pr = O.forward_project_view(g, np.ones((16, 16, 16)), 0)
    # theta = 0: y-slabs; the central cell's ray runs along -y through 16 voxels of 0.5 mm.
    assert pr[7, 10] == pytest.approx(16 * 0.5, rel=1e-14)


def test_uniform_box_x_branch_central_ray():
    g = O.Geometry(make_params(NumberOfImage=[12, 16, 16], NumberOfViews=4))
    pr = O.forward_project_view(g, np.ones((16, 16, 12)), 1)  # theta = pi/2: x-slabs
    assert pr[7, 10] == pytest.approx(12 * 0.5, rel=1e-12)


def test_uniform_box_off_centre_rays_closed_form():
    # A ray that stays inside the box crosses every slab; DD weights of one slab sum to 1 in
    # each direction, so p = n_slab * |d| * sqrt(1+s1^2) * sqrt(1+s2^2) (Q5 form of L).
    g = O.Geometry(make_params())
    pr = O.forward_project_view(g, np.ones((16, 16, 16)), 0)
    S = g.source_at(0)
    C, _ = O.detector_lattice(g, S, 0.0)
    s1 = (S[0] - C[0]) / (S[1] - C[1])
    s2 = (S[2] - C[2]) / (S[1] - C[1])
    cf = 16 * 0.5 * np.sqrt(1 + s1 ** 2) * np.sqrt(1 + s2 ** 2)
    np.testing.assert_allclose(pr[4:11, 6:15], cf[4:11, 6:15], rtol=1e-13)


# z-slab views: the detector rows sit about 1500 mm below the source height at 922 mm in-plane
# distance, so the rays are steep and every view away from the x and y axes selects the z-slab
# branch (python/GPUFuncs_pycuda.py:526-587); the volume sits where the rays cross at SAD.
ZSLAB = dict(NumberOfImage=[16, 16, 16], ImagePixelSpacing=[0.5, 0.5, 0.5], NumberOfDetectorPixels=[16, 16],
             DetectorPixelSize=[5.0, 5.0], SourceInit=[0, 100.0, 0.0], DetectorInit=[0, -100.0, -900.0],
             RotationOrigin=[0, 0, 0], DetectorOffset=[0.0, -1502.0], ReconCenter=[0.0, 0.0, -162.9],
             NumberOfViews=8, StartAngle=0.0, EndAngle=2 * np.pi)


def test_z_branch_selected_away_from_the_axes():
    g = O.Geometry(make_params(**ZSLAB))
    br = [O.select_branch(g.source_at(i), g.detector_at(i)) for i in range(g.nview)]
    assert br == [1, 2, 0, 2, 1, 2, 0, 2]


def test_z_branch_matches_brute_force():
    g = O.Geometry(make_params(**dict(ZSLAB, NumberOfImage=[7, 6, 8], NumberOfDetectorPixels=[9, 7])))
    vol = np.random.default_rng(11).random((8, 6, 7))
    for i in (1, 3):
        a = O.forward_project_view(g, vol, i)
        b = np.array(B.forward_view(g, vol, i))
        assert np.abs(a).max() > 0
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)


def test_z_branch_uniform_volume_closed_form():
    # A ray that stays inside the volume crosses every z slab; the footprint weights of one slab
    # sum to 1 in x and in y, so p = nz * |dz| * sqrt(1+s1^2) * sqrt(1+s2^2) with s1, s2 the
    # ray's x and y slopes against z (Q5 form of L).
    g = O.Geometry(make_params(**ZSLAB))
    xp, yp, zp = g.planes()
    for i in (1, 3, 5, 7):
        pr = O.forward_project_view(g, np.ones((16, 16, 16)), i)
        S = g.source_at(i)
        C, _ = O.detector_lattice(g, S, g.angles[i])
        s1 = (S[0] - C[0]) / (S[2] - C[2])
        s2 = (S[1] - C[1]) / (S[2] - C[2])
        cf = 16 * 0.5 * np.sqrt(1 + s1 ** 2) * np.sqrt(1 + s2 ** 2)
        # rays whose x and y stay a voxel inside the box from the top plane to the bottom one
        inside = np.ones(pr.shape, dtype=bool)
        for z in (zp[0], zp[-1]):
            x = S[0] + s1 * (z - S[2])
            y = S[1] + s2 * (z - S[2])
            inside &= (x > xp.min() + 1.0) & (x < xp.max() - 1.0) & (y > yp.min() + 1.0) & (y < yp.max() - 1.0)
        assert inside.sum() >= 4
        np.testing.assert_allclose(pr[inside], cf[inside], rtol=1e-12)


@pytest.mark.parametrize("shape", ["Flat", "Curved"])
def test_forward_matches_brute_force(shape):
    g = O.Geometry(make_params(NumberOfImage=[6, 7, 5], NumberOfDetectorPixels=[7, 5],
                               DetectorPixelSize=[2.0, 2.0], ImagePixelSpacing=[1.5, 1.5, 1.5],
                               NumberOfViews=5, DetectorShape=shape))
    vol = np.random.default_rng(3).random((5, 7, 6))
    for i in range(g.nview):
        a = O.forward_project_view(g, vol, i)
        b = np.array(B.forward_view(g, vol, i))
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)


@pytest.mark.parametrize("shape", ["Flat", "Curved"])
@pytest.mark.parametrize("mode", ["ref", "edge"])
def test_backproject_matches_brute_force(shape, mode):
    g = O.Geometry(make_params(NumberOfImage=[6, 7, 5], NumberOfDetectorPixels=[7, 5],
                               DetectorPixelSize=[2.0, 2.0], ImagePixelSpacing=[1.5, 1.5, 1.5],
                               NumberOfViews=5, DetectorShape=shape))
    q = np.random.default_rng(4).random((5, 7))
    for i in range(g.nview):
        a = O.backproject_view(g, q, i, mode)
        b = np.array(B.backproject_view(g, q, i, mode))
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)


def test_rotation_symmetry_between_x_and_y_branches():
    # Views theta and theta + pi/2 see the volume rotated by 90 degrees about z; the first
    # uses y-slabs, the second x-slabs, so a wrong axis or sign in either branch fails.
    g = O.Geometry(make_params(StartAngle=0.1, EndAngle=0.1 + 2 * np.pi, NumberOfViews=8,
                               NumberOfImage=[12, 12, 10]))
    vol = np.random.default_rng(5).random((10, 12, 12))
    # rotating the object by +90 deg (x,y) -> (-y, x) in world coordinates; with x increasing
    # and y decreasing along the array axes this is np.rot90 over (y, x) axes.
    rot = np.rot90(vol, k=1, axes=(1, 2))
    for i in (0, 1):
        a = O.forward_project_view(g, rot, i + 2)
        b = O.forward_project_view(g, vol, i)
        assert O.select_branch(g.source_at(i), g.detector_at(i)) != \
            O.select_branch(g.source_at(i + 2), g.detector_at(i + 2))
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * b.max())


@pytest.mark.parametrize("shape", ["Flat", "Curved"])
def test_forward_approximates_analytic_line_integrals(shape):
    g = O.Geometry(make_params(NumberOfImage=[40, 40, 40], NumberOfDetectorPixels=[48, 40],
                               DetectorPixelSize=[0.6, 0.6], ImagePixelSpacing=[0.5, 0.5, 0.5],
                               NumberOfViews=6, DetectorShape=shape))
    ell = [((2.0, -1.5, 1.0), (6.0, 4.0, 5.0), (0.3, 0.2, 0.1), 1.0),
           ((-3.0, 2.0, -2.0), (2.5, 3.0, 2.0), (1.0, 0.5, 0.0), 0.5)]
    xs, ys, zs = g.centres()
    vol = analytic.voxelise(ell, xs, ys, zs, sub=3)
    for i in (0, 1, 4):
        pr = O.forward_project_view(g, vol, i)
        S = g.source_at(i)
        C, _ = O.detector_lattice(g, S, g.angles[i])
        exact = analytic.line_integral(ell, S, np.moveaxis(C, 0, -1))
        if shape == "Curved":
            # The reference's curved lattice puts row k's edges at v_k - dv/2 and v_k - 3dv/2
            # (python/Reconstruction_pycuda.py:129), so row k integrates the row below it.
            pr, exact = pr[:-1], exact[1:]
        err = np.abs(pr - exact)
        # edges of the ellipsoids are steep, so bound the RMS error, the total and the
        # centre of mass (orientation of u and v) rather than the pointwise maximum
        assert np.sqrt(np.mean(err ** 2)) < 0.025 * exact.max()
        assert pr.sum() == pytest.approx(exact.sum(), rel=5e-3)
        for ax in (0, 1):
            idx = np.arange(pr.shape[1 - ax])
            com_dd = (pr.sum(ax) * idx).sum() / pr.sum()
            com_ex = (exact.sum(ax) * idx).sum() / exact.sum()
            assert abs(com_dd - com_ex) < 0.02


def test_adjoint_is_exact_transpose_of_forward():
    g = O.Geometry(make_params(NumberOfImage=[4, 5, 3], NumberOfDetectorPixels=[6, 4],
                               DetectorPixelSize=[1.5, 1.5], ImagePixelSpacing=[1.0, 1.0, 1.0],
                               NumberOfViews=3, StartAngle=0.3))
    n = 4 * 5 * 3
    A = np.zeros((3 * 4 * 6, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = O.forward_project(g, e.reshape(3, 5, 4)).ravel()
    y = np.random.default_rng(6).standard_normal((3, 4, 6))
    np.testing.assert_allclose(O.forward_adjoint(g, y).ravel(), A.T @ y.ravel(), rtol=0, atol=1e-12)


def test_backprojection_of_constant_data_closed_form():
    # Q = 1 everywhere: a voxel whose footprint lies inside the detector receives
    # sum_views dtheta * SAD^2 / (SAD - y_c)^2 (the overlap weights of one view sum to 1).
    g = O.Geometry(make_params(NumberOfImage=[8, 8, 6], NumberOfDetectorPixels=[40, 30], NumberOfViews=12))
    q = np.ones((12, 30, 40))
    vol = O.backproject(g, q, "ref")
    xs, ys, zs = g.centres()
    X, Y = np.meshgrid(xs, ys)
    expect = sum(g.dtheta * g.sad ** 2 / (g.sad - (Y * np.cos(t) - X * np.sin(t))) ** 2 for t in g.angles)
    np.testing.assert_allclose(vol, np.broadcast_to(expect, vol.shape), rtol=1e-12)


def test_ref_grid_is_edge_grid_shifted_half_pixel():
    # Q1: the reference's BP grid sits half a pixel off the true edges, which is the same as the
    # corrected grid with DetectorOffset = (du/2, -dv/2).
    kw = dict(NumberOfImage=[10, 9, 8], NumberOfDetectorPixels=[15, 11], NumberOfViews=4, StartAngle=0.2)
    g_ref = O.Geometry(make_params(**kw))
    g_edge = O.Geometry(make_params(DetectorOffset=[0.5, -0.5], **kw))
    q = np.random.default_rng(7).random((4, 11, 15))
    np.testing.assert_allclose(O.backproject(g_ref, q, "ref"), O.backproject(g_edge, q, "edge"),
                               rtol=0, atol=1e-12)


def test_backprojection_orientation_follows_the_ray():
    # A single hot cell (edge grid) lights voxels on the ray source -> cell centre:
    # at theta = 0, x = u*(SAD - y)/SDD and z = v*(SAD - y)/SDD.
    g = O.Geometry(make_params(NumberOfImage=[24, 24, 24], NumberOfDetectorPixels=[31, 31],
                               DetectorPixelSize=[0.75, 0.75], NumberOfViews=1, EndAngle=0.0))
    q = np.zeros((1, 31, 31))
    j, k = 22, 9  # u = (22-15)*0.75 = 5.25, v = (15-9)*0.75 = 4.5
    q[0, k, j] = 1.0
    vol = O.backproject_view(g, q[0], 0, "edge")
    xs, ys, zs = g.centres()
    for iy in (3, 12, 20):
        sl = vol[:, iy, :]
        iz, ix = np.unravel_index(np.argmax(sl), sl.shape)
        y = ys[iy]
        assert abs(xs[ix] - 5.25 * (1000 - y) / 1500) <= 0.5
        assert abs(zs[iz] - 4.5 * (1000 - y) / 1500) <= 0.5


@pytest.mark.parametrize("mode", ["ref", "edge"])
def test_fdk_of_uniform_sphere_recovers_density(mode):
    # End-to-end scale of Filtering() + backward(): FDK of analytic sphere projections.
    g = O.Geometry(make_params(NumberOfImage=[24, 24, 24], NumberOfDetectorPixels=[48, 40],
                               DetectorPixelSize=[0.8, 0.8], NumberOfViews=48, FilterType="ram-lak"))
    xs, ys, zs = g.centres()
    ell = [((0.0, 0.0, 0.0), (4.0, 4.0, 4.0), (0, 0, 0), 1.0)]
    proj = np.zeros((g.nview, g.nv, g.nu))
    for i in range(g.nview):
        S = g.source_at(i)
        C, _ = O.detector_lattice(g, S, g.angles[i])
        proj[i] = analytic.line_integral(ell, S, np.moveaxis(C, 0, -1))
    rec = O.fdk(g, proj.astype(np.float32), mode)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    assert abs(rec[r < 2.5].mean() - 1.0) < 0.02
    assert abs(rec[r > 5.5].mean()) < 0.02


def _blob_scan():
    p = make_params(NumberOfImage=[32, 32, 32], ImagePixelSpacing=[0.5, 0.5, 0.5],
                    NumberOfDetectorPixels=[48, 44], DetectorPixelSize=[0.8, 0.8], NumberOfViews=36,
                    FilterType="ram-lak")
    return O.Geometry(p), analytic.asymmetric_blobs(8.0)


@pytest.mark.parametrize("mode", ["edge", "ref"])
def test_fdk_of_asymmetric_object_puts_each_blob_in_place(mode):
    # Exact line integrals through the cell centres of the (separately pinned) lattice, filtered
    # and backprojected: each of three off-centre ellipsoids is reconstructed at its own centre.
    # This fixes the backprojector's u and v orientation and the filter's row handling against
    # the world frame.  The reference grid (Q1) sits half a cell off the true edges, which moves
    # the object by -dv/2 at the detector, -dv/2 * SAD/SDD at the isocentre, along z.
    g, ell = _blob_scan()
    proj = np.zeros((g.nview, g.nv, g.nu))
    for i in range(g.nview):
        S = g.source_at(i)
        C, _ = O.detector_lattice(g, S, g.angles[i])
        proj[i] = analytic.line_integral(ell, S, np.moveaxis(C, 0, -1))
    rec = O.fdk(g, proj.astype(np.float32), mode)
    xs, ys, zs = g.centres()
    err = analytic.centres_of_mass(rec, xs, ys, zs, ell) - np.array([e[0] for e in ell])
    if mode == "ref":
        err[:, 2] -= -0.5 * g.dv * g.sad / g.sdd
    assert np.abs(err[:, :2]).max() < 0.25 * g.dx
    assert np.abs(err[:, 2]).max() < 0.25 * g.dz if mode == "edge" else np.abs(err[:, 2]).max() < 0.1


def test_forward_projection_feeds_fdk_in_the_world_frame():
    # The DD forward projector composed with the pinned FDK chain: a flipped detector axis or
    # row order in the projector would mirror the blobs.
    g, ell = _blob_scan()
    xs, ys, zs = g.centres()
    vol = analytic.voxelise(ell, xs, ys, zs, sub=2)
    rec = O.fdk(g, O.forward_project(g, vol).astype(np.float32), "edge")
    err = analytic.centres_of_mass(rec, xs, ys, zs, ell) - np.array([e[0] for e in ell])
    assert np.abs(err).max() < 0.25 * g.dx
