"""Pins the C restatement of the oracle's FDK backprojection (oracle/c/dd_backproject.c) and the
batched filter of oracle/dd_c.py to the NumPy oracle they restate.

The C loop performs the oracle's FP64 operations in the oracle's order, so on a flat detector the
volumes must be identical bit for bit; on a curved detector libm's atan may differ from NumPy's
by an ulp.  The geometries cover both backprojection grids, helical source motion, detector and
reconstruction offsets, anisotropic voxels, a source close to the volume (wide footprints), a
fan-beam slice and weight zeroing.  A mistake planted in the NumPy oracle's overlap weights
must show.  The batched filter shares the oracle's spectrum and cosine weights and restates only
the FFT convolution; both round their FP64 rows to float32, so it is held to one float32 ulp of
the peak (two FFT libraries).
"""
import numpy as np
import pytest

from oracle import dd_c as C
from oracle import dd_oracle as O
from tests.conftest import make_params


@pytest.fixture(scope="module", autouse=True)
def built():
    import os
    if not os.path.exists(C.LIB_PATH):
        C.build()
    C.lib()


BASE = dict(StartAngle=0.1, EndAngle=2 * np.pi + 0.1, NumberOfDetectorPixels=[40, 32],
            NumberOfViews=12, NumberOfImage=[24, 22, 20], FilterType="shepp-logan")

CASES = {
    "flat": {},
    "helical": {"Pitch": 0.5},
    "offsets": {"DetectorOffset": [1.3, -0.7], "ReconCenter": [2.0, -1.0, 1.5]},
    "anisotropic": {"ImagePixelSpacing": [0.6, 0.4, 0.7], "DetectorPixelSize": [0.9, 1.2]},
    "near_source": {"SourceInit": [0, 60.0, 0], "DetectorInit": [0, -40.0, 0], "NumberOfImage": [16, 16, 32],
                    "ImagePixelSpacing": [1.0, 1.0, 1.0], "NumberOfDetectorPixels": [20, 40],
                    "DetectorPixelSize": [2.0, 2.0]},
    "fan_slice": {"NumberOfImage": [32, 32, 1], "NumberOfDetectorPixels": [64, 1], "FilterType": "ram-lak"},
    "curved": {"DetectorShape": "Curved"},
}


def _setup(name, seed=0):
    g = O.Geometry(make_params(**{**BASE, **CASES[name]}))
    q = np.random.default_rng(seed).standard_normal((g.nview, g.nv, g.nu)).astype(np.float32)
    return g, q


@pytest.mark.parametrize("mode", ["ref", "edge"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_c_backprojection_equals_numpy_oracle(name, mode):
    g, q = _setup(name)
    ref = O.backproject(g, q, mode)
    got = C.backproject(g, q, mode, threads=3)
    assert np.abs(ref).max() > 0
    if g.shape == "Curved":
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    else:
        assert np.array_equal(got, ref)


def test_c_backprojection_view_subset_thread_count_and_accumulation():
    g, q = _setup("offsets", seed=1)
    views = [3, 7, 10]
    ref = O.backproject(g, q, "ref", views=views)
    one = C.backproject(g, q[views], "ref", views=views, threads=1)
    many = C.backproject(g, q[views], "ref", views=views, threads=5)
    assert np.array_equal(one, ref) and np.array_equal(many, ref)
    acc = C.backproject(g, q[views[:1]], "ref", views=views[:1])
    C.backproject(g, q[views[1:]], "ref", views=views[1:], out=acc)
    assert np.array_equal(acc, ref)


def test_c_backprojection_weight_zeroing():
    g, q = _setup("anisotropic", seed=2)
    ref = O.backproject(g, q, "ref", zero_tol=0.05)
    assert np.array_equal(C.backproject(g, q, "ref", zero_tol=0.05), ref)
    assert not np.array_equal(ref, O.backproject(g, q, "ref"))


def test_c_backprojection_rejects_wrong_projection_shape():
    g, q = _setup("flat")
    with pytest.raises(ValueError):
        C.backproject(g, q[:, :, :-1], "ref")
    with pytest.raises(ValueError):
        C.backproject(g, q[:3], "ref", views=[0, 1])


@pytest.mark.parametrize("name", ["flat", "curved", "anisotropic", "fan_slice"])
def test_batched_filter_equals_numpy_oracle(name):
    g, q = _setup(name, seed=3)
    ref = O.filter_projections(g, q)
    got = C.filter_projections(g, q, threads=3, chunk=5)
    assert np.abs(got - ref).max() <= 2 ** -23 * np.abs(ref).max()


def test_planted_mistake_shows(monkeypatch):
    """The NumPy loop and the C loop are written independently: swapping the two end formulas
    of the v overlap in the NumPy oracle must break their agreement."""
    g, q = _setup("flat", seed=4)
    got_b = C.backproject(g, q, "ref", threads=2)
    orig = O._overlap_weights

    def swapped(c1, c2, s, e, k, planes, d, decreasing, zero_tol, dtype):
        return orig(c2, c1, s, e, k, planes, d, not decreasing, zero_tol, dtype) if decreasing else \
            orig(c1, c2, s, e, k, planes, d, decreasing, zero_tol, dtype)

    monkeypatch.setattr(O, "_overlap_weights", swapped)
    assert np.abs(O.backproject(g, q, "ref") - got_b).max() > 1e-3 * np.abs(got_b).max()


# ---------------------------------------------------------------- pixel-driven backprojection

from oracle import pixel_c as PC  # noqa: E402
from oracle import pixel_oracle as PO  # noqa: E402
from tests.conftest import make_cone_params  # noqa: E402

PIXEL_CASES = {
    "default": {},
    # odd detector and volume: the central column and slice sample exactly on a node at 0 degrees
    "nodes": {"DetectorWidth": 41, "DetectorHeight": 33, "ReconX": 25, "ReconY": 24, "ReconZ": 17,
              "NumberOfViews": 12},
    "short_scan": {"AngleCoverage": 1.2 * np.pi, "NumberOfViews": 20, "ReconZ": 9},
    # a volume wider than the detector's view: samples off the grid
    "off_grid": {"DetectorWidth": 20, "DetectorHeight": 12, "NumberOfViews": 10},
}


def _pixel_setup(name, seed=0):
    g = PO.PixelGeometry(make_cone_params(**PIXEL_CASES[name]))
    if name == "off_grid":
        g.x = g.x * 1.6
        g.z = g.z * 1.8
    proj = np.random.default_rng(seed).uniform(0.0, 2.0, (g.ns, g.nv, g.nu))
    return g, proj


@pytest.mark.parametrize("name", sorted(PIXEL_CASES))
def test_c_pixel_backprojection_and_margin_equal_numpy_oracle(name):
    g, proj = _pixel_setup(name)
    q = PO.filter_projections(g, proj)
    views = list(range(g.ns))
    ref = PO.backproject(g, q)
    got, margin = PC.backproject(g, q, views, threads=3, margin=True)
    assert np.abs(ref).max() > 0
    assert np.array_equal(got, ref)
    assert np.array_equal(margin, PO.edge_margin(g))
    if name == "off_grid":
        assert (margin > 0).all() and (ref == 0).any()


def test_c_pixel_backprojection_view_subset_and_scale():
    g, proj = _pixel_setup("short_scan", seed=1)
    q = PO.filter_projections(g, proj)
    views = [2, 9, 15]
    ref = PO.backproject(g, {i: q[i] for i in views}, scale=0.5, views=views)
    assert np.array_equal(PC.backproject(g, q[views], views, scale=0.5, threads=1), ref)


def test_c_pixel_planted_mistake_shows(monkeypatch):
    """Interpolating on nodes shifted by half a cell in the NumPy loop must break the equality."""
    g, proj = _pixel_setup("default", seed=2)
    q = PO.filter_projections(g, proj)
    got = PC.backproject(g, q, range(g.ns), threads=2)
    orig = PO.interpolate
    monkeypatch.setattr(PO, "interpolate", lambda q, ki, p, U, V: orig(q, ki + (ki[1] - ki[0]) / 2, p, U, V))
    assert np.abs(PO.backproject(g, q) - got).max() > 1e-3 * np.abs(got).max()


@pytest.mark.parametrize("name", ["default", "nodes"])
def test_batched_pixel_filter_equals_numpy_oracle(name):
    g, proj = _pixel_setup(name, seed=3)
    ref = PO.filter_projections(g, proj)
    got = PC.filter_projections(g, proj, threads=3, chunk=5)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


# ---------------------------------------------------------------- DD forward projection

@pytest.mark.parametrize("name", sorted(CASES))
def test_c_forward_equals_numpy_oracle(name):
    """Every view of each geometry (x- and y-slab views; the line tables come from the oracle,
    the slab loop is restated), bit for bit."""
    g, _ = _setup(name)
    vol = np.random.default_rng(5).random((g.nz, g.ny, g.nx)).astype(np.float32)
    got = C.forward_project(g, vol, threads=3)
    ref = O.forward_project(g, vol.astype(np.float64))
    assert np.abs(ref).max() > 0
    assert np.array_equal(got, ref)


def test_c_forward_z_slab_views_and_zero_tol():
    from tests.test_oracle_projectors import ZSLAB
    g = O.Geometry(make_params(**ZSLAB))
    branches = [O.fp_view_tables(g, i)[0] for i in range(g.nview)]
    assert 2 in branches and 0 in branches and 1 in branches
    vol = np.random.default_rng(6).random((g.nz, g.ny, g.nx)).astype(np.float32)
    for zt in (1e-6, None, 0.05):
        got = C.forward_project(g, vol, zero_tol=zt, threads=2)
        ref = O.forward_project(g, vol.astype(np.float64), zero_tol=zt)
        assert np.array_equal(got, ref), zt


def test_c_forward_planted_mistake_shows(monkeypatch):
    """Swapping the end formulas of the transverse overlap in the NumPy loop must break the
    equality (the C loop is written independently)."""
    g, _ = _setup("flat")
    vol = np.random.default_rng(7).random((g.nz, g.ny, g.nx)).astype(np.float32)
    got = C.forward_project_view(g, vol, 3, threads=2)
    orig = O._overlap_weights
    monkeypatch.setattr(O, "_overlap_weights",
                        lambda c1, c2, s, e, k, planes, d, dec, zt, dt: orig(c1, c2, s, e, k, planes, d, not dec, zt, dt))
    assert np.abs(O.forward_project_view(g, vol.astype(np.float64), 3) - got).max() > 1e-3 * np.abs(got).max()
