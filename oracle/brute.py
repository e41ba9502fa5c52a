"""Scalar pure-Python loops for tiny DD problems.

TEST INFRASTRUCTURE ONLY (same rule as ``dd_oracle``).  One Python loop
iteration per kernel thread, written the way the reference kernels are laid
out, so that the vectorised oracle's indexing (axis permutations, clipping,
masks) is checked against an independent scalar form:

* ``forward_view``: one "thread" per (slab, cell), python/GPUFuncs_pycuda.py:398-587;
* ``backproject_view``: one "thread" per voxel, python/GPUFuncs_pycuda.py:837-956.

The per-view tables come from ``dd_oracle.fp_view_tables`` (pinned separately
by closed forms), so only the kernel loops are restated here.
"""
from __future__ import annotations

import math

from . import dd_oracle as O


def _w(c1, c2, s, e, k, planes, d, zero_tol):
    lo, hi, width = min(c1, c2), max(c1, c2), abs(c1 - c2)
    dec = d < 0
    if s == e:
        w = 1.0
    elif k == s:
        w = (hi - planes[k + 1]) / width if dec else (planes[k + 1] - lo) / width
    elif k == e:
        w = (planes[k] - lo) / width if dec else (hi - planes[k]) / width
    else:
        w = abs(d) / width
    if zero_tol is not None and abs(w) < zero_tol:
        w = 0.0
    return w


def forward_view(g: O.Geometry, vol, i: int, zero_tol=1e-6):
    br, aU1, bU1, aU2, bU2, aV1, bV1, aV2, bV2, L = O.fp_view_tables(g, i)
    sa, ta, va = O.AXES[br]
    planes = g.planes()
    dvec = (g.dx, -g.dy, -g.dz)
    ps, pt, pv = planes[sa], planes[ta], planes[va]
    ds, dt, dv = dvec[sa], dvec[ta], dvec[va]
    nt, nvv = len(pt) - 1, len(pv) - 1
    out = [[0.0] * g.nu for _ in range(g.nv)]

    def voxel(n, k, l):
        idx = [0, 0, 0]
        idx[sa], idx[ta], idx[va] = n, k, l
        return float(vol[idx[2]][idx[1]][idx[0]])

    for n in range(len(ps) - 1):
        c = ps[n] + ds / 2
        for r in range(g.nv):
            for col in range(g.nu):
                t1 = aU1[r, col] * c + bU1[r, col]
                t2 = aU2[r, col] * c + bU2[r, col]
                v1 = aV1[r, col] * c + bV1[r, col]
                v2 = aV2[r, col] * c + bV2[r, col]
                i1, i2 = math.floor((t1 - pt[0]) / dt), math.floor((t2 - pt[0]) / dt)
                j1, j2 = math.floor((v1 - pv[0]) / dv), math.floor((v2 - pv[0]) / dv)
                st, et, sv, ev = min(i1, i2), max(i1, i2), min(j1, j2), max(j1, j2)
                for k in range(st, et + 1):
                    if not 0 <= k <= nt - 1:
                        continue
                    w1 = _w(t1, t2, st, et, k, pt, dt, zero_tol)
                    for l in range(sv, ev + 1):
                        if not 0 <= l <= nvv - 1:
                            continue
                        w2 = _w(v1, v2, sv, ev, l, pv, dv, zero_tol)
                        out[r][col] += voxel(n, k, l) * w1 * w2 * L[r, col]
    return out


def backproject_view(g: O.Geometry, q, i: int, mode="ref", zero_tol=None):
    ki, p, du, dv = O.bp_detector_grid(g, mode)
    th = g.angles[i]
    Sx, Sy = g.source[0], g.source[1]
    Sz = g.source_z0 + g.helical * th / (2 * math.pi)
    Dy = g.detector[1]
    R = g.sad
    c, s = math.cos(th), math.sin(th)
    xs, ys, zs = g.centres()
    hx, hy, hz = g.dx / 2, -g.dy / 2, -g.dz / 2
    out = [[[0.0] * g.nx for _ in range(g.ny)] for _ in range(g.nz)]
    for iz in range(g.nz):
        for iy in range(g.ny):
            for ix in range(g.nx):
                X, Y, Z = xs[ix], ys[iy], zs[iz]
                us, vs = [], []
                for sx in (1, -1):
                    for sy in (1, -1):
                        for sz in (1, -1):
                            xr = (X + sx * hx) * c + (Y + sy * hy) * s
                            yr = (Y + sy * hy) * c - (X + sx * hx) * s
                            us.append((Sx - xr) / (Sy - yr))
                            vs.append((Sz - (Z + sz * hz)) / (Sy - yr))
                if g.shape == "Curved":
                    cu1, cu2 = -math.atan(min(us)), -math.atan(max(us))
                else:
                    cu1 = min(us) * Dy + (Sx - min(us) * Sy)
                    cu2 = max(us) * Dy + (Sx - max(us) * Sy)
                cv1 = min(vs) * Dy + (Sz - min(vs) * Sy)
                cv2 = max(vs) * Dy + (Sz - max(vs) * Sy)
                i1, i2 = math.floor((cu1 - ki[0]) / du), math.floor((cu2 - ki[0]) / du)
                j1, j2 = math.floor((cv1 - p[0]) / dv), math.floor((cv2 - p[0]) / dv)
                su, eu, sv, ev = min(i1, i2), max(i1, i2), min(j1, j2), max(j1, j2)
                yc = Y * c - X * s
                W = R * R / ((R - yc) * (R - yc))
                acc = 0.0
                for k in range(sv, ev + 1):
                    if not 0 <= k <= g.nv - 1:
                        continue
                    w1 = _w(cv1, cv2, sv, ev, k, p, dv, zero_tol)
                    for l in range(su, eu + 1):
                        if not 0 <= l <= g.nu - 1:
                            continue
                        w2 = _w(cu1, cu2, su, eu, l, ki, du, zero_tol)
                        acc += float(q[k][l]) * w1 * w2 * W
                out[iz][iy][ix] = acc
    return out
