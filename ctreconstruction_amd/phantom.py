"""Seeded random-ellipsoid phantoms and Poisson projection noise, generated on the GPU.

The reference reads phantoms from raw files (python/FDK_512.py:19-22) that are not shipped;
these synthetic ones follow the benchmark description: K ellipsoids with centres
U(-0.8, 0.8)*FOV, semi-axes U(0.05, 0.4)*FOV, three Euler angles U(0, pi) and densities
U(-0.5, 1.0), summed, with FOV the half-extent of the volume.  Parameters are drawn with
``numpy.random.default_rng(seed)`` so a seed fixes the phantom on every machine; the voxel
values are evaluated at voxel centres on the device.
"""
from __future__ import annotations

import numpy as np
import torch

from .geometry import ScanGeometry

MU_WATER = 0.02  # mm^-1, multiplies densities before the Poisson step


def random_ellipsoids(n: int, fov, seed: int, ndim: int = 3):
    """List of (centre[3], semi_axes[3], euler[3], density) drawn from ``default_rng(seed)``.

    ``fov`` is the half-extent per axis (x, y, z); ``ndim=2`` draws ellipses in the x-y plane
    (z semi-axis infinite, single rotation angle)."""
    rng = np.random.default_rng(seed)
    fov = np.asarray(fov, dtype=np.float64)
    r = float(np.min(fov[:ndim]))
    out = []
    for _ in range(n):
        c = rng.uniform(-0.8, 0.8, size=3) * fov
        ax = rng.uniform(0.05, 0.4, size=3) * r
        ang = rng.uniform(0.0, np.pi, size=3)
        rho = rng.uniform(-0.5, 1.0)
        if ndim == 2:
            c[2] = 0.0
            ax[2] = np.inf
            ang[1:] = 0.0
        out.append((c, ax, ang, rho))
    return out


def _rotation(phi, theta, psi):
    cz, sz = np.cos(phi), np.sin(phi)
    cy, sy = np.cos(theta), np.sin(theta)
    cx, sx = np.cos(psi), np.sin(psi)
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    return Rz @ Ry @ Rx


def voxel_centres(geom: ScanGeometry):
    """Voxel centre coordinates (x, y, z) as in python/Reconstruction_pycuda.py:350-352."""
    rc = geom.recon_center
    x = rc[0] + (np.arange(geom.nx) - (geom.nx - 1) / 2.0) * geom.dx
    y = rc[1] + (np.arange(geom.ny) - (geom.ny - 1) / 2.0) * -geom.dy
    z = rc[2] + (np.arange(geom.nz) - (geom.nz - 1) / 2.0) * -geom.dz
    return x, y, z


def half_extent(geom: ScanGeometry):
    return np.array([geom.nx * geom.dx, geom.ny * geom.dy, geom.nz * geom.dz]) / 2.0


def ellipsoid_volume(geom: ScanGeometry, ellipsoids, device="cuda", slab: int = 32) -> torch.Tensor:
    """Sum of uniform ellipsoids sampled at voxel centres -> float32 [nz, ny, nx] on ``device``."""
    x, y, z = voxel_centres(geom)
    xt = torch.tensor(x, dtype=torch.float32, device=device)
    yt = torch.tensor(y, dtype=torch.float32, device=device)
    vol = torch.zeros(geom.vol_shape, dtype=torch.float32, device=device)
    for (c, ax, ang, rho) in ellipsoids:
        R = _rotation(*ang)
        inv2 = np.where(np.isinf(ax), 0.0, 1.0 / np.square(np.where(np.isinf(ax), 1.0, ax)))
        M = torch.tensor(R @ np.diag(inv2) @ R.T, dtype=torch.float32, device=device)
        cx, cy, cz = (float(v) for v in c)
        X = (xt - cx)[None, None, :]
        Y = (yt - cy)[None, :, None]
        for z0 in range(0, geom.nz, slab):
            zs = torch.tensor(z[z0:z0 + slab] - cz, dtype=torch.float32, device=device)[:, None, None]
            q = (M[0, 0] * X * X + M[1, 1] * Y * Y + M[2, 2] * zs * zs
                 + 2 * M[0, 1] * X * Y + 2 * M[0, 2] * X * zs + 2 * M[1, 2] * Y * zs)
            vol[z0:z0 + slab] += float(rho) * (q <= 1.0).to(torch.float32)
    return vol


def poisson_noise(proj: torch.Tensor, photons: float = 1e5, generator=None, chunk: int = 16) -> torch.Tensor:
    """``counts ~ Poisson(photons * exp(-p))``, returns ``-log(max(counts, 1) / photons)``.

    Processed ``chunk`` views at a time (the rate is formed in float64)."""
    out = torch.empty_like(proj, dtype=torch.float32)
    for v0 in range(0, proj.shape[0], chunk):
        lam = photons * torch.exp(-proj[v0:v0 + chunk].double())
        counts = torch.poisson(lam, generator=generator)
        out[v0:v0 + chunk] = -torch.log(torch.clamp(counts, min=1.0) / photons)
    return out
