"""FP64 NumPy oracle for the distance-driven (DD) projector pair and the FDK filter.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``ctreconstruction_amd``) never calls it.

It restates, in vectorised double precision, what the reference computes on
its pycuda path (the parity target):

* geometry: ``Reconstruction.__init__`` / angle grid / edge planes
  (python/Reconstruction_pycuda.py:64-99, 777-797, 349-354);
* detector lattice: ``FlatDetectorConstruction`` / ``CurvedDetectorConstruction``
  (python/Reconstruction_pycuda.py:113-168);
* forward projection: host tables in ``distance_forward``
  (python/Reconstruction_pycuda.py:826-1003) and kernels
  ``distance_project_on_{x,y,z}3`` (python/GPUFuncs_pycuda.py:398-587);
* filter: ``Filter`` / ``Filtering`` / ``filter_proj``
  (python/Reconstruction_pycuda.py:186-251);
* FDK backprojection: ``distance_backproj`` host (python/Reconstruction_pycuda.py:320-435)
  and kernel ``flat_distance_backproj_arb2`` (python/GPUFuncs_pycuda.py:837-956);
  the curved-detector variant follows the numba units
  (python/Reconstruction_numba.py:399-413, python/GPUFuncs_numba.py:537-685).

Every function says which reference lines it follows.  ``fp32=True`` casts
where the reference casts (float32 detector lattice, float32 slope/intercept
tables, float32 kernel arithmetic) so the reference's own single-precision
noise can be measured against the FP64 result.

Pinning (what checks each function independently of this file) lives in
``tests/test_oracle_*.py``: closed forms for the filter spectrum and for the
projection of a uniform box, brute-force loops on tiny inputs, the adjoint
identity, analytic ellipsoid line integrals and the FDK reconstruction of an
analytic phantom.
"""
from __future__ import annotations

import math

import numpy as np

PI = np.pi

FILTER_TYPES = ("ram-lak", "shepp-logan", "cosine", "hamming", "hann")

# Defaults of the reference's params dict (python/Reconstruction_pycuda.py:65-70).
DEFAULTS = {
    "SourceInit": [0, 0, 0], "DetectorInit": [0, 0, 0], "StartAngle": 0, "EndAngle": 0,
    "NumberOfDetectorPixels": [0, 0], "DetectorPixelSize": [0, 0], "NumberOfViews": 0,
    "ImagePixelSpacing": [0, 0, 0], "NumberOfImage": [0, 0, 0], "PhantomCenter": [0, 0, 0],
    "RotationOrigin": [0, 0, 0], "Method": "Distance", "FilterType": "ram-lak",
    "ReconCenter": [0, 0, 0], "cutoff": 1, "GPU": 0, "DetectorShape": "Flat", "Pitch": 0,
    "DetectorOffset": [0, 0],
}


class Geometry:
    """Scalars the reference derives from its params dict.

    Follows ``Reconstruction.__init__`` (python/Reconstruction_pycuda.py:64-99)
    and the angle grid (python/Reconstruction_pycuda.py:777-789).
    """

    def __init__(self, params: dict):
        p = dict(DEFAULTS)
        p.update(params)
        self.params = p
        self.nu, self.nv = (int(v) for v in p["NumberOfDetectorPixels"])
        self.du, self.dv = (float(v) for v in p["DetectorPixelSize"])
        self.nx, self.ny, self.nz = (int(v) for v in p["NumberOfImage"])
        self.dx, self.dy, self.dz = (float(v) for v in p["ImagePixelSpacing"])
        origin = np.asarray(p["RotationOrigin"], dtype=np.float64)
        self.source = np.asarray(p["SourceInit"], dtype=np.float64)
        self.detector = np.asarray(p["DetectorInit"], dtype=np.float64)
        self.sad = float(np.sqrt(np.sum((self.source - origin) ** 2.0)))
        self.sdd = float(np.sqrt(np.sum((self.source - self.detector) ** 2.0)))
        self.pitch = float(p["Pitch"])
        self.helical = self.pitch * (self.nv * self.dv) * self.sad / self.sdd
        self.nview = int(p["NumberOfViews"])
        self.angles = np.linspace(float(p["StartAngle"]), float(p["EndAngle"]), self.nview + 1)[:-1]
        if self.nview > 1:
            self.dtheta = float(self.angles[1] - self.angles[0])
        else:
            self.dtheta = float(p["EndAngle"]) - float(p["StartAngle"])
        self.recon_center = np.asarray(p["ReconCenter"], dtype=np.float64)
        self.det_offset = np.asarray(p.get("DetectorOffset", [0, 0]), dtype=np.float64)
        self.shape = p["DetectorShape"]
        self.filter_type = p["FilterType"]
        self.cutoff = float(p["cutoff"])
        self.source_z0 = float(self.source[2])
        self.detector_z0 = float(self.detector[2])

    # -- per-view source/detector (python/Reconstruction_pycuda.py:828-830)
    def source_at(self, i: int) -> np.ndarray:
        th = self.angles[i]
        return np.array([-self.sad * np.sin(th), self.sad * np.cos(th),
                         self.source_z0 + self.helical * th / (2 * PI)])

    def detector_at(self, i: int) -> np.ndarray:
        th = self.angles[i]
        return np.array([(self.sdd - self.sad) * np.sin(th), -(self.sdd - self.sad) * np.cos(th),
                         self.detector_z0 + self.helical * th / (2 * PI)])

    # -- edge planes of the forward projector (python/Reconstruction_pycuda.py:792-797)
    def planes(self):
        dx, dy, dz = self.dx, -self.dy, -self.dz
        rc = self.recon_center
        xp = rc[0] + (np.arange(self.nx + 1) - (self.nx - 1) / 2.0) * dx - dx / 2
        yp = rc[1] + (np.arange(self.ny + 1) - (self.ny - 1) / 2.0) * dy - dy / 2
        zp = rc[2] + (np.arange(self.nz + 1) - (self.nz - 1) / 2.0) * dz - dz / 2
        return xp, yp, zp

    # -- voxel centres of the backprojector (python/Reconstruction_pycuda.py:350-352)
    def centres(self):
        dx, dy, dz = self.dx, -self.dy, -self.dz
        rc = self.recon_center
        xc = rc[0] + (np.arange(self.nx) - (self.nx - 1) / 2.0) * dx
        yc = rc[1] + (np.arange(self.ny) - (self.ny - 1) / 2.0) * dy
        zc = rc[2] + (np.arange(self.nz) - (self.nz - 1) / 2.0) * dz
        return xc, yc, zc


def select_branch(S: np.ndarray, D: np.ndarray) -> int:
    """Slab axis of one view: 0 = x-slabs, 1 = y-slabs, 2 = z-slabs.

    Same comparisons, in the same order and precision, as
    python/Reconstruction_pycuda.py:852-853 and :921-922 (ties go to x, then y).
    """
    ax, ay, az = abs(S[0] - D[0]), abs(S[1] - D[1]), abs(S[2] - D[2])
    if ax >= ay and ax >= az:
        return 0
    if ay >= ax and ay >= az:
        return 1
    return 2


def detector_lattice(g: Geometry, S: np.ndarray, theta: float, fp32: bool = False):
    """Cell centres C[3,nv,nu] and cell-edge lattice B[3,nv+1,nu+1].

    Flat: python/Reconstruction_pycuda.py:139-168.  Curved: :113-137 (its edge
    rows use ``- dv/2``, one pixel off the flat lattice's ``+ dv/2``; kept).
    The reference stores both arrays as float32; ``fp32`` reproduces that.
    """
    eu = np.array([np.cos(theta), np.sin(theta), 0.0])
    ew = np.array([np.sin(theta), -np.cos(theta), 0.0])
    nu, nv = g.nu, g.nv
    if g.shape == "Flat":
        u = (np.arange(nu) - (nu - 1.0) / 2.0) * g.du + g.det_offset[0]
        v = (np.arange(nv) - (nv - 1.0) / 2.0) * -1.0 * g.dv + g.det_offset[1]
        u2 = (np.arange(nu + 1) - (nu - 1) / 2.0) * g.du - g.du / 2.0 + g.det_offset[0]
        v2 = (np.arange(nv + 1) - (nv - 1) / 2.0) * -1.0 * g.dv + g.dv / 2.0 + g.det_offset[1]
        U, V = np.meshgrid(u, v)
        U2, V2 = np.meshgrid(u2, v2)
        C = np.stack([S[k] + U * eu[k] + g.sdd * ew[k] + V * (k == 2) for k in range(3)])
        B = np.stack([S[k] + U2 * eu[k] + g.sdd * ew[k] + V2 * (k == 2) for k in range(3)])
    elif g.shape == "Curved":
        da = g.du / g.sdd
        u = (np.arange(nu) - (nu - 1) / 2.0) * da + g.det_offset[0]
        v = (np.arange(nv) - (nv - 1) / 2.0) * -1.0 * g.dv + g.det_offset[1]
        u2 = (np.arange(nu + 1) - (nu - 1) / 2.0) * da - da / 2.0 + g.det_offset[0]
        v2 = (np.arange(nv + 1) - (nv - 1) / 2.0) * -1.0 * g.dv - g.dv / 2.0 + g.det_offset[1]
        U, V = np.meshgrid(u, v)
        U2, V2 = np.meshgrid(u2, v2)
        C = np.stack([S[k] + g.sdd * np.sin(U) * eu[k] + g.sdd * np.cos(U) * ew[k] + V * (k == 2)
                      for k in range(3)])
        B = np.stack([S[k] + g.sdd * np.sin(U2) * eu[k] + g.sdd * np.cos(U2) * ew[k] + V2 * (k == 2)
                      for k in range(3)])
    else:
        raise ValueError("DetectorShape must be 'Flat' or 'Curved'")
    if fp32:
        C = C.astype(np.float32)
        B = B.astype(np.float32)
    return C, B


# Axis roles per branch: (slab axis, U-transverse axis, V-transverse axis).
# y-branch: python/GPUFuncs_pycuda.py:398-461; x-branch :463-524; z-branch :526-587.
AXES = {0: (0, 1, 2), 1: (1, 0, 2), 2: (2, 0, 1)}


def fp_view_tables(g: Geometry, i: int, fp32: bool = False):
    """Per-view branch + footprint line tables (python/Reconstruction_pycuda.py:828-864, 924-935, 993-1003).

    Returns (branch, aU1, bU1, aU2, bU2, aV1, bV1, aV2, bV2, L) with the lines
    ``t = a*s + b`` giving a transverse coordinate ``t`` at slab coordinate ``s``.
    """
    S = g.source_at(i)
    D = g.detector_at(i)
    C, B = detector_lattice(g, S, g.angles[i], fp32)
    br = select_branch(S, D)
    sa, ta, va = AXES[br]
    # boundary points of the four footprint lines (python/Reconstruction_pycuda.py:840-846)
    U1 = [B[0, :-1, :-1], B[1, :-1, :-1], C[2]]
    U2 = [B[0, 1:, 1:], B[1, 1:, 1:], C[2]]
    V1 = [C[0], C[1], B[2, 1:, 1:]]
    V2 = [C[0], C[1], B[2, :-1, :-1]]
    f8 = lambda a: np.asarray(a, dtype=np.float64)  # noqa: E731  (float32 lattice -> FP64 math, as numpy does)
    aU1 = (S[ta] - f8(U1[ta])) / (S[sa] - f8(U1[sa]))
    bU1 = -aU1 * S[sa] + S[ta]
    aU2 = (S[ta] - f8(U2[ta])) / (S[sa] - f8(U2[sa]))
    bU2 = -aU2 * S[sa] + S[ta]
    aV1 = (S[va] - f8(V1[va])) / (S[sa] - f8(V1[sa]))
    bV1 = -aV1 * S[sa] + S[va]
    aV2 = (S[va] - f8(V2[va])) / (S[sa] - f8(V2[sa]))
    bV2 = -aV2 * S[sa] + S[va]
    s1 = (S[ta] - f8(C[ta])) / (S[sa] - f8(C[sa]))
    s2 = (S[va] - f8(C[va])) / (S[sa] - f8(C[sa]))
    d = (g.dx, -g.dy, -g.dz)[sa]
    L = abs(d) / (np.cos(np.arctan(s1)) * np.cos(np.arctan(s2)))
    tabs = [aU1, bU1, aU2, bU2, aV1, bV1, aV2, bV2, L]
    if fp32:
        tabs = [t.astype(np.float32) for t in tabs]
    return (br, *tabs)


def _overlap_weights(c1, c2, s_idx, e_idx, k, planes, d, decreasing, zero_tol, dtype):
    """DD overlap weight of cell ``k`` for the footprint [min(c1,c2), max(c1,c2)].

    Case structure of python/GPUFuncs_pycuda.py:427-438 (increasing planes) and
    :441-452 (decreasing planes); weights below ``zero_tol`` are zeroed.
    """
    lo = np.minimum(c1, c2)
    hi = np.maximum(c1, c2)
    width = np.abs(c1 - c2)
    kc = np.clip(k, 0, len(planes) - 2)
    pk = planes[kc]
    pk1 = planes[kc + 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        if decreasing:
            w_s = (hi - pk1) / width
            w_e = (pk - lo) / width
        else:
            w_s = (pk1 - lo) / width
            w_e = (hi - pk) / width
        w_i = dtype(abs(d)) / width
    w = np.where(s_idx == e_idx, dtype(1.0), np.where(k == s_idx, w_s, np.where(k == e_idx, w_e, w_i)))
    if zero_tol is not None:
        w = np.where(np.abs(w) < zero_tol, dtype(0.0), w)
    return w.astype(dtype)


def forward_project_view(g: Geometry, vol: np.ndarray, i: int, fp32: bool = False,
                         zero_tol: float | None = 1e-6) -> np.ndarray:
    """One view of the DD forward projection -> [nv, nu].

    Kernel semantics of ``distance_project_on_{x,y,z}3`` (python/GPUFuncs_pycuda.py:398-587):
    for every detector cell and every slab, the footprint [t1,t2] x [v1,v2] of
    the cell's four boundary lines at the slab centre is overlapped with the
    voxel grid; the sum of voxel * w1 * w2 over slabs is scaled by the cell's
    intersection length L.
    """
    dt = np.float32 if fp32 else np.float64
    br, aU1, bU1, aU2, bU2, aV1, bV1, aV2, bV2, L = fp_view_tables(g, i, fp32)
    sa, ta, va = AXES[br]
    planes = [np.asarray(p, dtype=dt) for p in g.planes()]
    dvec = [dt(g.dx), dt(-g.dy), dt(-g.dz)]
    ps, pt, pv = planes[sa], planes[ta], planes[va]
    ds, dtt, dvv = dvec[sa], dvec[ta], dvec[va]
    nt, nvv = len(pt) - 1, len(pv) - 1
    # volume viewed as [slab, t, v]
    vol_ax = np.transpose(np.asarray(vol, dtype=dt), (2, 1, 0))  # -> [x, y, z]
    vol_stv = np.transpose(vol_ax, (sa, ta, va))
    out = np.zeros((g.nv, g.nu), dtype=np.float64)
    for n in range(len(ps) - 1):
        c = dt(ps[n] + ds / dt(2))
        t1 = (aU1 * c + bU1).astype(dt)
        t2 = (aU2 * c + bU2).astype(dt)
        v1 = (aV1 * c + bV1).astype(dt)
        v2 = (aV2 * c + bV2).astype(dt)
        it1 = np.floor((t1 - pt[0]) / dtt).astype(np.int64)
        it2 = np.floor((t2 - pt[0]) / dtt).astype(np.int64)
        iv1 = np.floor((v1 - pv[0]) / dvv).astype(np.int64)
        iv2 = np.floor((v2 - pv[0]) / dvv).astype(np.int64)
        st, et = np.minimum(it1, it2), np.maximum(it1, it2)
        sv, ev = np.minimum(iv1, iv2), np.maximum(iv1, iv2)
        wt_max = int(np.max(et - st)) + 1
        wv_max = int(np.max(ev - sv)) + 1
        acc = np.zeros((g.nv, g.nu), dtype=dt)
        slab = vol_stv[n]
        for a in range(wt_max):
            k = st + a
            ok_k = (k <= et) & (k >= 0) & (k <= nt - 1)
            if not ok_k.any():
                continue
            w1 = _overlap_weights(t1, t2, st, et, k, pt, dtt, dtt < 0, zero_tol, dt)
            for b in range(wv_max):
                l = sv + b
                ok = ok_k & (l <= ev) & (l >= 0) & (l <= nvv - 1)
                if not ok.any():
                    continue
                w2 = _overlap_weights(v1, v2, sv, ev, l, pv, dvv, dvv < 0, zero_tol, dt)
                f = slab[np.clip(k, 0, nt - 1), np.clip(l, 0, nvv - 1)]
                term = (f * w1 * w2 * L).astype(dt)
                acc = acc + np.where(ok, term, dt(0))
        out += acc
    return out


def forward_project(g: Geometry, vol: np.ndarray, fp32: bool = False,
                    zero_tol: float | None = 1e-6, views=None) -> np.ndarray:
    """All views -> proj[nView, nv, nu] (python/Reconstruction_pycuda.py:762-1059)."""
    views = range(g.nview) if views is None else views
    views = list(views)
    out = np.zeros((len(views), g.nv, g.nu), dtype=np.float64)
    for n, i in enumerate(views):
        out[n] = forward_project_view(g, vol, i, fp32, zero_tol)
    return out


def forward_adjoint(g: Geometry, proj: np.ndarray, zero_tol: float | None = 1e-6) -> np.ndarray:
    """Exact transpose A^T of :func:`forward_project` (scatter form of the same weights).

    No reference counterpart: the reference backprojects with the FDK-weighted
    operator, not with A^T.  Used to pin the matched backprojector.
    """
    vol_x = np.zeros((g.nx, g.ny, g.nz), dtype=np.float64)
    planes = g.planes()
    dvec = [g.dx, -g.dy, -g.dz]
    for i in range(g.nview):
        br, aU1, bU1, aU2, bU2, aV1, bV1, aV2, bV2, L = fp_view_tables(g, i)
        sa, ta, va = AXES[br]
        ps, pt, pv = planes[sa], planes[ta], planes[va]
        ds, dtt, dvv = dvec[sa], dvec[ta], dvec[va]
        nt, nvv = len(pt) - 1, len(pv) - 1
        view_vol = np.transpose(vol_x, (sa, ta, va))  # a view into vol_x
        y = proj[i]
        for n in range(len(ps) - 1):
            c = ps[n] + ds / 2
            t1, t2 = aU1 * c + bU1, aU2 * c + bU2
            v1, v2 = aV1 * c + bV1, aV2 * c + bV2
            it1 = np.floor((t1 - pt[0]) / dtt).astype(np.int64)
            it2 = np.floor((t2 - pt[0]) / dtt).astype(np.int64)
            iv1 = np.floor((v1 - pv[0]) / dvv).astype(np.int64)
            iv2 = np.floor((v2 - pv[0]) / dvv).astype(np.int64)
            st, et = np.minimum(it1, it2), np.maximum(it1, it2)
            sv, ev = np.minimum(iv1, iv2), np.maximum(iv1, iv2)
            slab = view_vol[n]
            for a in range(int(np.max(et - st)) + 1):
                k = st + a
                ok_k = (k <= et) & (k >= 0) & (k <= nt - 1)
                if not ok_k.any():
                    continue
                w1 = _overlap_weights(t1, t2, st, et, k, pt, dtt, dtt < 0, zero_tol, np.float64)
                for b in range(int(np.max(ev - sv)) + 1):
                    l = sv + b
                    ok = ok_k & (l <= ev) & (l >= 0) & (l <= nvv - 1)
                    if not ok.any():
                        continue
                    w2 = _overlap_weights(v1, v2, sv, ev, l, pv, dvv, dvv < 0, zero_tol, np.float64)
                    np.add.at(slab, (k[ok], l[ok]), (y * w1 * w2 * L)[ok])
    return np.transpose(vol_x, (2, 1, 0)).copy()


# ----------------------------------------------------------------------------- filter

def zero_padded_length(nu: int) -> int:
    """``2**ceil(log2(2*(nu-1)))`` (python/Reconstruction_pycuda.py:236)."""
    return int(2 ** (math.ceil(math.log2(2.0 * (nu - 1)))))


def ramp_spectrum(N: int, tau: float, filter_type: str, cutoff: float) -> np.ndarray:
    """Centred filter magnitude of length N-1 (python/Reconstruction_pycuda.py:186-221).

    Discrete Ram-Lak taps h(0)=1/(8 tau^2), h(odd n)=-1/(2 pi^2 tau^2 n^2);
    |fftshift(fft(h))|, then the window and the frequency mask
    ``|w| > pi*cutoff/(2 tau)`` (w in rad/sample, tau in mm, as the reference).
    Shepp-Logan's DC bin is set to ``H[DC]*sin(0) = 0`` as the reference does.
    """
    if filter_type not in FILTER_TYPES:
        raise ValueError(f"unknown FilterType {filter_type!r}")
    x = np.arange(0, N) - (N - 1) / 2.0
    h = np.zeros(len(x))
    h[x == 0] = 1 / (8 * tau ** 2)
    odd = np.where(x % 2.0 == 1)
    h[odd] = -0.5 / (PI * tau * x[odd]) ** 2
    h = h[:-1]
    H = np.abs(np.fft.fftshift(np.fft.fft(h)))
    w = 2 * PI * x[:-1] / (N - 1)
    if filter_type == "shepp-logan":
        zero = w == 0
        arg = w / (2.0 * cutoff)
        with np.errstate(divide="ignore", invalid="ignore"):
            sinc = np.where(zero, 0.0, np.sin(arg) / np.where(zero, 1.0, arg))
        H = H * sinc
    elif filter_type == "cosine":
        H = H * np.cos(w / (2.0 * cutoff))
    elif filter_type == "hamming":
        H = H * (0.54 + 0.46 * np.cos(w / cutoff))
    elif filter_type == "hann":
        H = H * (0.5 + 0.5 * np.cos(w / cutoff))
    H[np.abs(w) > PI * cutoff / (2.0 * tau)] = 0
    return H


def fdk_weights(g: Geometry) -> np.ndarray:
    """Cosine pre-weight [nv, nu] at the reference's cell-edge coordinates (python/Reconstruction_pycuda.py:224-240)."""
    ki = (np.arange(0, g.nu + 1) - g.nu / 2.0) * g.du + g.det_offset[0]
    p = (np.arange(0, g.nv + 1) - g.nv / 2.0) * -1.0 * g.dv + g.det_offset[1]
    R = g.sad
    D = g.sdd - R
    kk, pp = np.meshgrid(ki[0:-1] * R / (R + D), p[0:-1] * R / (R + D))
    return R / np.sqrt(R ** 2.0 + kk ** 2.0 + pp ** 2.0)


def filter_projections(g: Geometry, proj: np.ndarray) -> np.ndarray:
    """``Filtering()``: cosine weight, zero-padded FFT convolution, crop, * tau.

    python/Reconstruction_pycuda.py:223-251; rows are filtered independently.
    Output is float32, as the reference stores it.
    """
    nu = g.nu
    zp = zero_padded_length(nu)
    R = g.sad
    D = g.sdd - R
    tau = g.du * R / (R + D)
    H = ramp_spectrum(zp + 1, tau, g.filter_type, g.cutoff)
    Hs = np.fft.ifftshift(H)
    W = fdk_weights(g)
    out = np.empty(proj.shape, dtype=np.float32)
    for i in range(proj.shape[0]):
        wp = W * np.asarray(proj[i], dtype=np.float64)
        q = np.real(np.fft.ifft(Hs * np.fft.fft(wp, zp, axis=-1), axis=-1))
        out[i] = (q[:, :nu] * tau).astype(np.float32)
    return out


# ----------------------------------------------------------------------------- backprojection

def bp_detector_grid(g: Geometry, mode: str = "ref"):
    """Detector cell boundaries (u_plane, v_plane, du, dv) used by the backprojector.

    ``ref``: the reference grid, half a pixel off the forward projector's edges
    (python/Reconstruction_pycuda.py:353-354; curved units from
    python/Reconstruction_numba.py:399-413).  ``edge``: the forward projector's
    true cell edges (the reference's own commented-out grid, :356-357).
    """
    du = g.du
    dv = -g.dv
    if g.shape == "Curved":
        du = du / g.sdd
    if mode == "ref":
        ki = (np.arange(0, g.nu + 1) - (g.nu - 1) / 2.0) * du
        p = (np.arange(0, g.nv + 1) - (g.nv - 1) / 2.0) * dv
    elif mode == "edge":
        ki = (np.arange(0, g.nu + 1) - g.nu / 2.0) * du
        p = (np.arange(0, g.nv + 1) - g.nv / 2.0) * dv
    else:
        raise ValueError(mode)
    return ki + g.det_offset[0], p + g.det_offset[1], du, dv


def backproject_view(g: Geometry, q: np.ndarray, i: int, mode: str = "ref",
                     zero_tol: float | None = None, fp32: bool = False) -> np.ndarray:
    """One view of the FDK-weighted DD backprojection -> [nz, ny, nx] (not yet * dtheta).

    Kernel semantics of ``flat_distance_backproj_arb2`` (python/GPUFuncs_pycuda.py:837-956)
    and, for a curved detector, ``curved_distance_backproj_arb``
    (python/GPUFuncs_numba.py:537-685): rotate the 8 voxel corners by the view
    angle into the source frame, take min/max of the u- and v-slopes through
    the (unrotated) source, intersect with the detector plane y = DetectorInit_y,
    overlap with the detector grid and weight by SAD^2/(SAD - y_c)^2.
    """
    dt = np.float32 if fp32 else np.float64
    ki, p, du, dv = bp_detector_grid(g, mode)
    ki, p, du, dv = ki.astype(dt), p.astype(dt), dt(du), dt(dv)
    th = g.angles[i]
    S = g.source.copy()
    S[2] = g.source_z0 + g.helical * th / (2 * PI)
    Sx, Sy, Sz = (dt(s) for s in S)
    Dy = dt(g.detector[1])
    R = dt(g.sad)
    c, s = (np.cos(dt(th)), np.sin(dt(th)))
    xc, yc_, zc = (a.astype(dt) for a in g.centres())
    X = xc[None, None, :]
    Y = yc_[None, :, None]
    Z = zc[:, None, None]
    hx, hy, hz = dt(g.dx / 2.0), dt(-g.dy / 2.0), dt(-g.dz / 2.0)
    y_rot = Y * c - X * s
    us, vs = [], []
    for sx in (1, -1):
        for sy in (1, -1):
            xr = (X + sx * hx) * c + (Y + sy * hy) * s
            yr = (Y + sy * hy) * c - (X + sx * hx) * s
            us.append((Sx - xr) / (Sy - yr))
            for sz in (1, -1):
                vs.append((Sz - (Z + sz * hz)) / (Sy - yr))
    shape = (g.nz, g.ny, g.nx)
    # the u-footprint depends on the voxel column only ([1, ny, nx]); v on the voxel ([nz, ny, nx])
    u_min = np.minimum.reduce(us)
    u_max = np.maximum.reduce(us)
    v_min = np.broadcast_to(np.minimum.reduce(vs), shape)
    v_max = np.broadcast_to(np.maximum.reduce(vs), shape)
    if g.shape == "Curved":
        cu1, cu2 = -np.arctan(u_min), -np.arctan(u_max)
    else:
        cu1 = u_min * Dy + (Sx - u_min * Sy)
        cu2 = u_max * Dy + (Sx - u_max * Sy)
    cv1 = v_min * Dy + (Sz - v_min * Sy)
    cv2 = v_max * Dy + (Sz - v_max * Sy)
    iu1 = np.floor((cu1 - ki[0]) / du).astype(np.int64)
    iu2 = np.floor((cu2 - ki[0]) / du).astype(np.int64)
    iv1 = np.floor((cv1 - p[0]) / dv).astype(np.int64)
    iv2 = np.floor((cv2 - p[0]) / dv).astype(np.int64)
    su, eu = np.minimum(iu1, iu2), np.maximum(iu1, iu2)
    sv, ev = np.minimum(iv1, iv2), np.maximum(iv1, iv2)
    W = (R * R) / ((R - y_rot) * (R - y_rot))  # [1, ny, nx]
    acc = np.zeros(shape, dtype=dt)
    u_taps = []
    for b in range(int(np.max(eu - su)) + 1):
        l = su + b
        ok_l = (l <= eu) & (l >= 0) & (l <= g.nu - 1)
        if ok_l.any():
            w2 = _overlap_weights(cu1, cu2, su, eu, l, ki, du, False, zero_tol, dt)
            u_taps.append((np.clip(l, 0, g.nu - 1), np.where(ok_l, w2 * W, dt(0)).astype(dt)))
    for a in range(int(np.max(ev - sv)) + 1):
        k = sv + a
        ok_k = (k <= ev) & (k >= 0) & (k <= g.nv - 1)
        if not ok_k.any():
            continue
        w1 = np.where(ok_k, _overlap_weights(cv1, cv2, sv, ev, k, p, dv, True, zero_tol, dt), dt(0))
        kc = np.clip(k, 0, g.nv - 1)
        row_sum = np.zeros(shape, dtype=dt)
        for lc, w2W in u_taps:
            row_sum = row_sum + (q[kc, lc].astype(dt) * w2W).astype(dt)
        acc = acc + (row_sum * w1).astype(dt)
    return acc


def backproject(g: Geometry, q: np.ndarray, mode: str = "ref", zero_tol: float | None = None,
                fp32: bool = False, views=None) -> np.ndarray:
    """``distance_backproj`` (python/Reconstruction_pycuda.py:320-435): sum over views of
    the per-view backprojection of ``Q = proj * dtheta``."""
    views = range(g.nview) if views is None else views
    vol = np.zeros((g.nz, g.ny, g.nx), dtype=np.float64)
    for i in views:
        qi = np.asarray(q[i], dtype=np.float64) * g.dtheta
        if fp32:
            qi = qi.astype(np.float32)
        vol += backproject_view(g, qi, i, mode, zero_tol, fp32)
    return vol


def fdk(g: Geometry, proj: np.ndarray, mode: str = "ref") -> np.ndarray:
    """Filtering() followed by backward() (python/FDK_512.py:31-33)."""
    return backproject(g, filter_projections(g, proj), mode)
