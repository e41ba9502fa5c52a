"""Pins the oracle's geometry conventions (reference rows a1-a4 and the FDK pre-weight) by
closed forms that do not reuse its formulas.

Reference: ``Reconstruction.__init__`` (python/Reconstruction_pycuda.py:64-99), the angle grid
(:788-789), the edge planes (:792-797) and voxel centres (:350-352), the per-view source and
detector (:828-830), ``FlatDetectorConstruction`` / ``CurvedDetectorConstruction`` (:113-168)
and the cosine pre-weight of ``filter_proj`` (:224-240).

The checks are written as functions of the lattice and weight builders so that the same checks
also run against builders with a planted mistake (flipped e_u, rows counted bottom-up, the
detector at the wrong distance, a missing magnification): each such mutant must fail at least
one check.  Expected values are hand-evaluated numbers or independent constructions (the
theta = 0 detector rotated about z; the ray from the source through a cell-edge point
intersected with the isocentre plane).
"""
import math

import numpy as np
import pytest

from oracle import dd_oracle as O
from tests.conftest import make_params

# a small detector with offsets, so that centre/edge and offset sign mistakes show
LAT = dict(NumberOfDetectorPixels=[5, 3], DetectorPixelSize=[2.0, 3.0], DetectorOffset=[0.5, -0.25])


def _geom(**kw):
    p = dict(LAT)
    p.update(kw)
    return O.Geometry(make_params(**p))


# ----------------------------------------------------------------------------- a1 / a2 / a3

def test_scalars_from_the_params_dict():
    g = O.Geometry(make_params(SourceInit=[0, 1000.0, 0], DetectorInit=[0, -500.0, 0],
                               RotationOrigin=[3.0, 4.0, 0.0], Pitch=0.5,
                               NumberOfDetectorPixels=[10, 8], DetectorPixelSize=[1.0, 1.5]))
    assert g.sad == pytest.approx(math.hypot(3.0, 996.0), rel=1e-15)
    assert g.sdd == 1500.0
    # H = Pitch * nv * dv * SAD / SDD
    assert g.helical == pytest.approx(0.5 * 8 * 1.5 * math.hypot(3.0, 996.0) / 1500.0, rel=1e-15)


def test_angle_grid_drops_the_end_point():
    g = O.Geometry(make_params(StartAngle=0.0, EndAngle=2 * np.pi, NumberOfViews=8))
    np.testing.assert_allclose(g.angles, np.arange(8) * np.pi / 4, rtol=0, atol=1e-15)
    assert g.dtheta == pytest.approx(np.pi / 4, rel=1e-15)


def test_source_and_detector_centre_at_0_and_90_degrees():
    # SAD = |SourceInit - RotationOrigin| counts z as well: the origin shares the source height
    g = O.Geometry(make_params(NumberOfViews=4, SourceInit=[0, 1000.0, 2.0], DetectorInit=[0, -500.0, 2.0],
                               RotationOrigin=[0, 0, 2.0]))
    np.testing.assert_allclose(g.source_at(0), [0.0, 1000.0, 2.0], atol=1e-12)
    np.testing.assert_allclose(g.detector_at(0), [0.0, -500.0, 2.0], atol=1e-12)
    # theta = pi/2: the source has moved to -x, the detector to +x
    np.testing.assert_allclose(g.source_at(1), [-1000.0, 0.0, 2.0], atol=1e-12)
    np.testing.assert_allclose(g.detector_at(1), [500.0, 0.0, 2.0], atol=1e-12)


def test_helical_source_height():
    g = O.Geometry(make_params(NumberOfViews=4, EndAngle=4 * np.pi, Pitch=1.0,
                               NumberOfDetectorPixels=[10, 6], DetectorPixelSize=[1.0, 1.0]))
    # H = 1 * 6 * 1 * 1000 / 1500 = 4 mm per turn; view 1 sits at theta = pi
    assert g.source_at(1)[2] == pytest.approx(2.0, abs=1e-12)
    assert g.source_at(2)[2] == pytest.approx(4.0, abs=1e-12)


def test_edge_planes_and_voxel_centres():
    g = O.Geometry(make_params(NumberOfImage=[4, 3, 2], ImagePixelSpacing=[0.5, 1.0, 2.0],
                               ReconCenter=[1.0, -2.0, 3.0]))
    xp, yp, zp = g.planes()
    # x increases from ReconCenter_x - nx dx / 2; y and z decrease from RC + n d / 2
    np.testing.assert_allclose(xp, [0.0, 0.5, 1.0, 1.5, 2.0], atol=1e-15)
    np.testing.assert_allclose(yp, [-0.5, -1.5, -2.5, -3.5], atol=1e-15)
    np.testing.assert_allclose(zp, [5.0, 3.0, 1.0], atol=1e-15)
    xc, yc, zc = g.centres()
    np.testing.assert_allclose(xc, [0.25, 0.75, 1.25, 1.75], atol=1e-15)
    np.testing.assert_allclose(yc, [-1.0, -2.0, -3.0], atol=1e-15)
    np.testing.assert_allclose(zc, [4.0, 2.0], atol=1e-15)


# ----------------------------------------------------------------------------- a4 detector lattice

def check_flat_lattice(lattice, tol=1e-9):
    """Closed forms of the flat lattice.  ``lattice(g, S, theta)`` returns (C, B); ``tol`` is the
    position tolerance in mm (2e-4 for a lattice stored in float32, as the reference stores it)."""
    g = _geom()
    S0 = np.array([0.0, 1000.0, 0.0])
    C, B = lattice(g, S0, 0.0)
    assert C.shape == (3, 3, 5) and B.shape == (3, 4, 6)
    # theta = 0, du = 2, dv = 3, offsets (0.5, -0.25): u_j = (j - 2) 2 + 0.5, v_k = (1 - k) 3 - 0.25;
    # every cell sits on the plane y = -(SDD - SAD) = -500
    np.testing.assert_allclose(C[:, 0, 0], [-3.5, -500.0, 2.75], atol=tol)
    np.testing.assert_allclose(C[:, 2, 4], [4.5, -500.0, -3.25], atol=tol)
    np.testing.assert_allclose(C[:, 1, 2], [0.5, -500.0, -0.25], atol=tol)
    u = (np.arange(5) - 2.0) * 2.0 + 0.5
    v = (1.0 - np.arange(3)) * 3.0 - 0.25
    np.testing.assert_allclose(C[0], np.broadcast_to(u, (3, 5)), atol=tol)
    np.testing.assert_allclose(C[2], np.broadcast_to(v[:, None], (3, 5)), atol=tol)
    np.testing.assert_allclose(C[1], -500.0, atol=tol)
    # edges: u - du/2 on the left, v + dv/2 on top (row 0 is the top row)
    np.testing.assert_allclose(B[:, 0, 0], [-4.5, -500.0, 4.25], atol=tol)
    np.testing.assert_allclose(B[:, 3, 5], [5.5, -500.0, -4.75], atol=tol)
    np.testing.assert_allclose(B[0, 0], np.arange(6) * 2.0 - 4.5, atol=tol)
    np.testing.assert_allclose(B[2, :, 0], 4.25 - 3.0 * np.arange(4), atol=tol)
    # theta = pi/2: the source sits at (-1000, 0, 0); the detector plane is x = +500, u runs along +y
    S1 = np.array([-1000.0, 0.0, 0.0])
    C1, B1 = lattice(g, S1, np.pi / 2)
    np.testing.assert_allclose(C1[:, 0, 0], [500.0, -3.5, 2.75], atol=tol)
    np.testing.assert_allclose(C1[:, 2, 4], [500.0, 4.5, -3.25], atol=tol)
    np.testing.assert_allclose(B1[:, 0, 0], [500.0, -4.5, 4.25], atol=tol)
    # any theta: the theta = 0 lattice rotated about z by theta
    th = 0.7
    c, s = math.cos(th), math.sin(th)
    rot = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    Ct, Bt = lattice(g, rot @ S0, th)
    np.testing.assert_allclose(Ct, np.einsum("ij,jkl->ikl", rot, C), atol=tol)
    np.testing.assert_allclose(Bt, np.einsum("ij,jkl->ikl", rot, B), atol=tol)


def check_curved_lattice(lattice, tol=1e-9):
    """The equiangular arc: centred on the source with radius SDD, cell j at angle
    (j - (nu-1)/2) du/SDD + off_u from the central ray, edge rows at v - dv/2."""
    g = _geom(DetectorShape="Curved", DetectorOffset=[0.001, -0.25])
    S0 = np.array([0.0, 1000.0, 0.0])
    C, B = lattice(g, S0, 0.0)
    r = np.hypot(C[0] - S0[0], C[1] - S0[1])
    np.testing.assert_allclose(r, 1500.0, rtol=0, atol=tol)
    ang = np.arctan2(C[0] - S0[0], S0[1] - C[1])  # angle from the central ray (towards -y)
    np.testing.assert_allclose(ang[0], (np.arange(5) - 2.0) * 2.0 / 1500.0 + 0.001, atol=tol / 1000.0)
    np.testing.assert_allclose(C[2][:, 0], (1.0 - np.arange(3)) * 3.0 - 0.25, atol=tol)
    # the curved lattice's edge rows sit at v_k - dv/2 (python/Reconstruction_pycuda.py:129)
    np.testing.assert_allclose(B[2][:, 0], (1.0 - np.arange(4)) * 3.0 - 1.5 - 0.25, atol=tol)
    angb = np.arctan2(B[0] - S0[0], S0[1] - B[1])
    np.testing.assert_allclose(angb[0], (np.arange(6) - 2.5) * 2.0 / 1500.0 + 0.001, atol=tol / 1000.0)
    np.testing.assert_allclose(np.hypot(B[0] - S0[0], B[1] - S0[1]), 1500.0, rtol=0, atol=tol)
    S1 = np.array([-1000.0, 0.0, 0.0])
    C1, _ = lattice(g, S1, np.pi / 2)
    np.testing.assert_allclose(C1[0, :, 2], 500.0 - 1500.0 * (1 - math.cos(0.001)), atol=tol)
    np.testing.assert_allclose(C1[1, 1], 1500.0 * np.sin((np.arange(5) - 2.0) * 2.0 / 1500.0 + 0.001), atol=tol)


def oracle_lattice(g, S, th):
    return O.detector_lattice(g, S, th)


def test_flat_lattice_closed_form():
    check_flat_lattice(oracle_lattice)


def test_curved_lattice_closed_form():
    check_curved_lattice(oracle_lattice)


def test_lattice_float32_storage():
    # the reference stores C and B as float32 (:154, :163)
    g = _geom()
    C, B = O.detector_lattice(g, g.source_at(0), 0.0, fp32=True)
    assert C.dtype == np.float32 and B.dtype == np.float32


def _flip_eu(g, S, th):
    C, B = O.detector_lattice(g, S, th)
    # a lattice whose u axis points the other way: mirror about the central column
    return C[:, :, ::-1].copy(), B[:, :, ::-1].copy()


def _rows_bottom_up(g, S, th):
    C, B = O.detector_lattice(g, S, th)
    return C[:, ::-1].copy(), B[:, ::-1].copy()


def _detector_at_sad(g, S, th):
    # the detector plane through the isocentre instead of at SDD from the source
    h = O.Geometry(dict(g.params, DetectorInit=[0, 0.0, 0]))
    h.det_offset = g.det_offset
    return O.detector_lattice(h, S, th)


def _centres_for_edges(g, S, th):
    C, B = O.detector_lattice(g, S, th)
    # edges at +du/2 instead of -du/2
    sh = np.zeros(3)
    sh[:2] = [math.cos(th) * g.du, math.sin(th) * g.du]
    return C, B + sh[:, None, None]


@pytest.mark.parametrize("mutant", [_flip_eu, _rows_bottom_up, _detector_at_sad, _centres_for_edges])
def test_planted_lattice_mistakes_fail_the_checks(mutant):
    with pytest.raises(AssertionError):
        check_flat_lattice(mutant)


# ----------------------------------------------------------------------------- FDK pre-weight

def check_fdk_weights(weights):
    """W[k, j] = SAD / |S - P_kj|: the cosine of the ray from the source through the top-left
    cell-edge point B[:, k, j] of the theta = 0 lattice, with P_kj that ray's intersection
    with the isocentre plane y = 0 (python/Reconstruction_pycuda.py:224-240)."""
    g = _geom(NumberOfDetectorPixels=[6, 4], DetectorPixelSize=[30.0, 45.0], DetectorOffset=[12.0, -9.0])
    W = weights(g)
    assert W.shape == (4, 6)
    S = np.array([0.0, 1000.0, 0.0])
    _, B = O.detector_lattice(g, S, 0.0)
    t = S[1] / (S[1] - B[1, :-1, :-1])  # ray parameter of the plane y = 0
    P = S[:, None, None] + t * (B[:, :-1, :-1] - S[:, None, None])
    np.testing.assert_allclose(W, 1000.0 / np.linalg.norm(P - S[:, None, None], axis=0), rtol=1e-13)
    # hand-evaluated: cell (0, 0) has the edge point u = -3*30 + 12 = -78, v = 2*45 - 9 = 81 at
    # the detector; scaled by SAD/SDD = 2/3: (-52, 54)
    assert W[0, 0] == pytest.approx(1000.0 / math.sqrt(1000.0 ** 2 + 52.0 ** 2 + 54.0 ** 2), rel=1e-14)


def test_fdk_weights_follow_the_ray_through_the_cell_edge():
    check_fdk_weights(O.fdk_weights)


def _weights_without_magnification(g):
    h = O.Geometry(dict(g.params, DetectorInit=[0, 0.0, 0]))  # SDD = SAD: R/(R+D) = 1
    h.det_offset = g.det_offset
    return O.fdk_weights(h)


def _weights_at_cell_centres(g):
    h = O.Geometry(dict(g.params))
    h.det_offset = g.det_offset + np.array([g.du / 2, -g.dv / 2])
    return O.fdk_weights(h)


def _weights_rows_flipped(g):
    return O.fdk_weights(g)[::-1].copy()


@pytest.mark.parametrize("mutant", [_weights_without_magnification, _weights_at_cell_centres,
                                    _weights_rows_flipped])
def test_planted_weight_mistakes_fail_the_checks(mutant):
    with pytest.raises(AssertionError):
        check_fdk_weights(mutant)
