"""Iterative reconstruction on the DD projector pair: SART and ASD-POCS.

Two backprojectors are available:

* ``"ref"`` — the reference's own loop semantics (python/SART.py:29-59,
  python/ASD_POCS.py:130-202): the FDK-weighted DD backprojector (``backward()``, which scales by
  dtheta and by SAD^2/(SAD - y_c)^2), normalisations ``norm1 = BP(1)`` and ``norm2 = FP(1)``,
  and the reference's placement of eps.
* ``"matched"`` — the update is the gradient of the weighted least-squares data term
  ``0.5 * sum W (A x - p)^2`` with W = 1/(A 1 + eps), taken by torch autograd through
  :class:`DDProjectorFunction`, whose backward is the matched backprojector A^T; it is scaled
  by 1/(A^T 1 + eps) (the SART normalisation).

``tv_gradient`` is the normalised total-variation gradient of python/ASD_POCS.py:55-77;
``tv_gradient_autograd`` obtains the same vector by differentiating the TV functional.
All functions accept one volume [nz, ny, nx] (projections [nView, nv, nu]) or a batch.
"""
from __future__ import annotations

import weakref

import torch

from . import projector as P
from .autograd import DDProjectorFunction, adjoint_batch, forward_batch
from .geometry import ScanGeometry


def _bp_ref(proj: torch.Tensor, geom: ScanGeometry, mode: str = "ref") -> torch.Tensor:
    """Reference ``backward()``: FDK-weighted DD backprojection times dtheta (no filter)."""
    def one(p, g, out=None):
        return P.dd_backproject(p, g, mode=mode, out=out)
    if proj.dim() == 3:
        return one(proj.contiguous(), geom)
    out = torch.empty((proj.shape[0],) + geom.vol_shape, dtype=torch.float32, device=proj.device)
    for b in range(proj.shape[0]):
        one(proj[b].contiguous(), geom, out=out[b])
    return out


# per geometry object: {(backprojector, device): (A 1, B 1)}, freed with the geometry
_NORM_CACHE: "weakref.WeakKeyDictionary[ScanGeometry, dict]" = weakref.WeakKeyDictionary()


def normalisers(geom: ScanGeometry, backprojector: str, device, eps: float):
    """(row sums A 1, column sums B 1) for the chosen backprojector B; cached per geometry object
    (held weakly: the cache goes when the geometry does)."""
    per_geom = _NORM_CACHE.setdefault(geom, {})
    key = (backprojector, str(device))
    hit = per_geom.get(key)
    if hit is not None:
        return hit
    ones_v = torch.ones(geom.vol_shape, device=device)
    ones_p = torch.ones(geom.proj_shape, device=device)
    a1 = P.dd_forward(ones_v, geom)
    if backprojector == "matched":
        b1 = P.dd_adjoint(ones_p, geom)
    elif backprojector == "edge":
        b1 = P.dd_backproject(ones_p, geom, mode="edge")
    else:
        b1 = _bp_ref(ones_p, geom)
    per_geom[key] = (a1, b1)
    return a1, b1


def sart_reference(proj0: torch.Tensor, geom: ScanGeometry, n_iter: int, alpha: float = 1.0,
                   eps: float = 1e-5, x0: torch.Tensor | None = None, callback=None) -> torch.Tensor:
    """The reference's SART loop (python/SART.py:29-59), including its update
    ``x <- x - alpha * (BP((FP x - p)/(FP 1 + eps)) / BP 1 + eps)`` (eps outside the division)."""
    geom.check_supported()
    a1, b1 = normalisers(geom, "ref", proj0.device, eps)
    shape = (proj0.shape[0],) + geom.vol_shape if proj0.dim() == 4 else geom.vol_shape
    x = torch.zeros(shape, device=proj0.device) if x0 is None else x0.clone()
    for it in range(n_iter):
        resid = (forward_batch(x, geom) - proj0) / (a1 + eps)
        x = x - alpha * (_bp_ref(resid, geom) / b1 + eps)
        if callback is not None:
            callback(it, x)
    return x


def sart(proj0: torch.Tensor, geom: ScanGeometry, n_iter: int, relaxation: float = 0.5,
         x0: torch.Tensor | None = None, nonneg: bool = False, eps: float = 1e-6,
         backprojector: str = "matched", callback=None) -> torch.Tensor:
    """SART with relaxation lambda: x <- x + lambda * B((p - A x) / (A 1 + eps)) / (B 1 + eps).

    With ``backprojector="matched"`` the step direction comes from torch autograd of the
    weighted data term (the backward pass runs the matched backprojector A^T); with ``"ref"``
    B is the reference's FDK-weighted backprojector.
    """
    geom.check_supported()
    a1, b1 = normalisers(geom, backprojector, proj0.device, eps)
    # rays that miss the volume carry no information (their measured value is noise); they are
    # left out as in python/ASD_POCS.py:175, instead of being amplified by 1 / eps
    w = torch.where(a1 > 0, 1.0 / (a1 + eps), torch.zeros_like(a1))
    x = (torch.zeros((proj0.shape[0],) + geom.vol_shape if proj0.dim() == 4 else geom.vol_shape,
                     device=proj0.device) if x0 is None else x0.detach().clone())
    for it in range(n_iter):
        if backprojector == "matched":
            xv = x.detach().requires_grad_(True)
            r = DDProjectorFunction.apply(xv, geom) - proj0
            loss = 0.5 * (w * r * r).sum()
            (grad,) = torch.autograd.grad(loss, xv)
            step = -grad
        else:
            step = _bp_ref((proj0 - forward_batch(x, geom)) * w, geom, "edge" if backprojector == "edge" else "ref")
        x = x.detach() + relaxation * step / (b1 + eps)
        if nonneg:
            x = torch.clamp_min(x, 0.0)
        if callback is not None:
            callback(it, x)
    return x.detach()


def _backward_differences(x: torch.Tensor):
    """Backward differences along the last three axes, at every voxel that has a predecessor
    along all three (the voxel (i, j, k) of ``x`` with i, j, k >= 1 sits at (i-1, j-1, k-1))."""
    core = x[..., 1:, 1:, 1:]
    return (core - x[..., :-1, 1:, 1:], core - x[..., 1:, :-1, 1:], core - x[..., 1:, 1:, :-1])


def _tv_terms(img: torch.Tensor, eps: float):
    """Gradient of TV(x) = sum_v sqrt(eps + |D x(v)|^2) (D: backward differences) at the interior
    voxels c, before normalisation: the voxel's own term sum_k D_k x(c) / |D x(c)|_eps minus,
    for each axis k, D_k x(c + e_k) / |D x(c + e_k)|_eps from its successor along k.  This is the
    vector of python/ASD_POCS.py:55-72."""
    d = _backward_differences(img)
    norm = torch.sqrt(eps + d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    own = ((d[0] + d[1] + d[2]) / norm)[..., :-1, :-1, :-1]
    ratio = [dk / norm for dk in d]
    succ = ratio[0][..., 1:, :-1, :-1] + ratio[1][..., :-1, 1:, :-1] + ratio[2][..., :-1, :-1, 1:]
    return own - succ


def _normalise_pad(dtv: torch.Tensor, eps: float) -> torch.Tensor:
    dims = tuple(range(dtv.dim() - 3, dtv.dim()))
    norm = torch.sqrt(torch.sum(dtv ** 2, dim=dims, keepdim=True))
    norm = torch.where(norm.abs() < eps, torch.ones_like(norm), norm)
    return torch.nn.functional.pad(dtv / norm, (1, 1, 1, 1, 1, 1))


def tv_gradient(img: torch.Tensor, eps: float = 1e-7) -> torch.Tensor:
    """Normalised TV gradient with zero border (python/ASD_POCS.py:55-77)."""
    return _normalise_pad(_tv_terms(img, eps), eps)


def tv_gradient_autograd(img: torch.Tensor, eps: float = 1e-7) -> torch.Tensor:
    """The same vector as :func:`tv_gradient`, from autograd of
    TV(x) = sum over voxels v with all backward neighbours of sqrt(eps + |grad_b x(v)|^2),
    restricted to interior voxels and normalised as the reference does."""
    x = img.detach().requires_grad_(True)
    d = _backward_differences(x)
    tv = torch.sqrt(eps + d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    # the reference's sum covers the interior voxels and their +1 neighbours in one axis
    (grad,) = torch.autograd.grad(tv.sum(), x)
    inner = grad[..., 1:-1, 1:-1, 1:-1]
    return _normalise_pad(inner, eps)


def asd_pocs(proj0: torch.Tensor, geom: ScanGeometry, n_outer: int = 20, na: int = 5, ng: int = 20,
             x0: torch.Tensor | None = None, alpha: float = 0.007, beta: float = 1.0,
             beta_red: float = 0.999, r_max: float = 0.99, alpha_red: float = 0.95, eps: float = 1e-7,
             backprojector: str = "ref", tv: str = "reference", callback=None) -> torch.Tensor:
    """ASD-POCS (python/ASD_POCS.py:130-202): ``na`` non-negative SART steps, then ``ng``
    normalised TV steepest-descent steps of size alpha * |dx_POCS|, with the reference's
    adaptive rules for alpha and beta.  ``tv="autograd"`` takes the TV gradient by autograd.
    ``x0=None`` starts from the FDK reconstruction, as the reference script does.  As in the
    reference, the TV steps are not followed by a non-negativity projection."""
    geom.check_supported()
    a1, b1 = normalisers(geom, "matched" if backprojector == "matched" else "ref", proj0.device, eps)
    bp = (lambda r: adjoint_batch(r, geom)) if backprojector == "matched" else (lambda r: _bp_ref(r, geom))
    tvg = tv_gradient_autograd if tv == "autograd" else tv_gradient
    if x0 is None:
        # the reference starts from the FDK reconstruction (python/ASD_POCS.py:142-143)
        x = (torch.stack([P.fdk(p.contiguous(), geom) for p in proj0]) if proj0.dim() == 4
             else P.fdk(proj0.contiguous(), geom))
    else:
        x = x0.detach().clone()
    # a batch holds independent reconstructions: distances and alpha are per volume
    dims = tuple(range(x.dim() - 3, x.dim()))
    dist = lambda a, b: torch.sqrt(torch.sum((a - b) ** 2, dim=dims, keepdim=True))  # noqa: E731
    alpha = torch.full_like(x[..., :1, :1, :1], alpha)
    for it in range(n_outer):
        f0 = x.clone()
        for _ in range(na):
            r = (proj0 - forward_batch(x, geom)) / (a1 + eps)
            r = torch.where(a1 == 0, torch.zeros_like(r), r)
            upd = bp(r)
            upd = torch.where(b1 == 0, torch.zeros_like(upd), upd)
            x = torch.clamp_min(x + beta * upd / (b1 + eps), 0.0)
        dp = dist(x, f0)
        dtvg = alpha * dp
        f0 = x.clone()
        for _ in range(ng):
            x = x - dtvg * tvg(x, eps)
        dg = dist(x, f0)
        alpha = torch.where(dg > r_max * dp, alpha * alpha_red, alpha)
        beta *= beta_red
        if callback is not None:
            callback(it, x)
    return x
