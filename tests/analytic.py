"""Analytic ellipsoid phantoms for tests: exact voxel-independent line integrals.

A uniform ellipsoid {x : (x-c)^T M (x-c) <= 1} with density rho has, along the
line S + t d, the chord length 2*sqrt(B^2 - A*C)/A * |d| (A = d^T M d,
B = d^T M (S-c), C = (S-c)^T M (S-c) - 1).  These closed forms check the
oracle's forward projector and FDK chain independently of its own formulas.
"""
from __future__ import annotations

import numpy as np


def rotation(phi, theta, psi):
    cz, sz = np.cos(phi), np.sin(phi)
    cy, sy = np.cos(theta), np.sin(theta)
    cx, sx = np.cos(psi), np.sin(psi)
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    return Rz @ Ry @ Rx


def quad_matrix(axes, angles):
    R = rotation(*angles)
    return R @ np.diag(1.0 / np.asarray(axes, float) ** 2) @ R.T


def line_integral(ellipsoids, S, P):
    """Sum over ellipsoids (centre, axes, angles, rho) of rho * chord(S->P); P may be [...,3]."""
    P = np.asarray(P, float)
    d = P - S
    out = np.zeros(P.shape[:-1])
    for (c, axes, ang, rho) in ellipsoids:
        M = quad_matrix(axes, ang)
        o = S - np.asarray(c, float)
        A = np.einsum("...i,ij,...j->...", d, M, d)
        B = np.einsum("...i,ij,j->...", d, M, o)
        C = o @ M @ o - 1.0
        disc = np.maximum(B * B - A * C, 0.0)
        out += rho * 2.0 * np.sqrt(disc) / A * np.linalg.norm(d, axis=-1)
    return out


def voxelise(ellipsoids, xc, yc, zc, sub=1):
    """Volume [nz, ny, nx] of summed densities at voxel centres (sub x sub x sub supersampling)."""
    dx = xc[1] - xc[0] if len(xc) > 1 else 1.0
    dy = yc[1] - yc[0] if len(yc) > 1 else 1.0
    dz = zc[1] - zc[0] if len(zc) > 1 else 1.0
    offs = (np.arange(sub) + 0.5) / sub - 0.5
    vol = np.zeros((len(zc), len(yc), len(xc)))
    for oz in offs:
        for oy in offs:
            for ox in offs:
                Z, Y, X = np.meshgrid(zc + oz * dz, yc + oy * dy, xc + ox * dx, indexing="ij")
                pts = np.stack([X, Y, Z], axis=-1)
                for (c, axes, ang, rho) in ellipsoids:
                    M = quad_matrix(axes, ang)
                    r = pts - np.asarray(c, float)
                    inside = np.einsum("...i,ij,...j->...", r, M, r) <= 1.0
                    vol += rho * inside
    return vol / sub ** 3


def asymmetric_blobs(half):
    """Three separated ellipsoids placed off-centre and without mirror symmetry in a cube of
    half-extent ``half`` (mm): (centre, semi-axes, Euler angles, density)."""
    h = float(half)
    return [((0.40 * h, -0.30 * h, 0.35 * h), (0.20 * h, 0.14 * h, 0.16 * h), (0.3, 0.0, 0.0), 1.0),
            ((-0.45 * h, 0.35 * h, -0.20 * h), (0.14 * h, 0.22 * h, 0.15 * h), (0.0, 0.4, 0.0), 0.7),
            ((-0.10 * h, -0.50 * h, -0.45 * h), (0.16 * h, 0.12 * h, 0.12 * h), (0.0, 0.0, 0.5), 0.5)]


def centres_of_mass(rec, xs, ys, zs, ellipsoids, grow=1.6):
    """Centre of mass (x, y, z) of ``rec`` [nz, ny, nx] inside a box ``grow`` times each
    ellipsoid's bounding box, one row per ellipsoid."""
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    out = []
    for (c, axes, ang, rho) in ellipsoids:
        r = grow * float(np.max(axes))
        box = (np.abs(X - c[0]) <= r) & (np.abs(Y - c[1]) <= r) & (np.abs(Z - c[2]) <= r)
        w = np.where(box, rec, 0.0)
        m = w.sum()
        out.append([(w * X).sum() / m, (w * Y).sum() / m, (w * Z).sum() / m])
    return np.array(out)
