"""NumPy restatement of the reference's iterative loops on the FP64 DD oracle.

TEST INFRASTRUCTURE ONLY (same rule as ``dd_oracle``).

* ``sart_reference`` follows python/SART.py:29-59: FP / FDK-weighted BP (``backward()``
  without filtering) and ``x <- x - alpha * (BP((FP x - p)/(FP 1 + eps)) / BP 1 + eps)``.
* ``tv_grad`` follows python/ASD_POCS.py:55-77 (normalised TV gradient, zero border).

Pinned in ``tests/test_oracle_iterative.py``: ``tv_grad`` against central finite differences
of the TV functional; ``sart_reference`` against one hand-evaluated update on a 1-voxel /
1-cell system and against its fixed point (consistent data).
"""
from __future__ import annotations

import numpy as np

from . import dd_oracle as O


def sart_reference(g: O.Geometry, proj0: np.ndarray, n_iter: int, alpha: float = 1.0,
                   eps: float = 1e-5) -> np.ndarray:
    norm1 = O.backproject(g, np.ones((g.nview, g.nv, g.nu)), "ref")
    norm2 = O.forward_project(g, np.ones((g.nz, g.ny, g.nx)))
    x = np.zeros((g.nz, g.ny, g.nx))
    for _ in range(n_iter):
        r = (O.forward_project(g, x) - proj0) / (norm2 + eps)
        x = x - alpha * (O.backproject(g, r, "ref") / norm1 + eps)
    return x


def tv_grad(image: np.ndarray, eps: float = 1e-7) -> np.ndarray:
    im = np.asarray(image, dtype=np.float64)
    c = im[1:-1, 1:-1, 1:-1]
    a = c - im[0:-2, 1:-1, 1:-1]
    b = c - im[1:-1, 0:-2, 1:-1]
    d = c - im[1:-1, 1:-1, 0:-2]
    out = (a + b + d) / np.sqrt(eps + a ** 2 + b ** 2 + d ** 2)
    for ax in range(3):
        nb = [slice(1, -1)] * 3
        nb[ax] = slice(2, None)
        n = im[tuple(nb)]
        diffs = []
        for bx in range(3):
            sl = list(nb)
            if bx == ax:
                sl[bx] = slice(1, -1)
            else:
                sl[bx] = slice(0, -2)
            diffs.append(n - im[tuple(sl)])
        own = diffs[ax]
        out = out - own / np.sqrt(eps + sum(dd ** 2 for dd in diffs))
    norm = np.sqrt(np.sum(out ** 2))
    if abs(norm) < eps:
        norm = 1.0
    return np.pad(out / norm, 1)
