"""Learned prior + SART on sparse views: pair generation, training, inference and evaluation.

Scan (the sparse-view benchmark): 256^3 volume at 0.5 mm of seeded random ellipsoids times
mu_water = 0.02/mm, 384 x 384 flat detector at 1.0 mm, SID 1000, SDD 1500, Poisson noise at
1e5 photons.  A training pair is made by the method itself from one phantom:

* input  = 12-iteration SART (relaxation 0.5, matched backprojector) of the 60-view noisy scan,
* target = FDK of the 720-view noisy scan (Shepp-Logan window).

The network (``prior.ResidualPrior``) sees 3 adjacent slices of the input and predicts the
target's centre slice.  At inference the learned reconstruction is: 12 SART iterations, the
network on every slice, then 5 SART data-consistency iterations starting from its output (the
weights' metadata records both counts; ``pipeline_for`` reads them).
Plain SART is the fallback and the arbiter: if the learned result fits the measured projections
worse than the SART iterate it started from (projection-data residual), plain SART continues
instead.  Reconstructions are in mm^-1; the network works in units of mu_water (x 50).
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import iterative as IT
from . import phantom
from . import prior as PR
from . import projector as P
from .geometry import ScanGeometry

SCALE = 1.0 / phantom.MU_WATER  # network units


def scan_params(views: int, n: int = 256, voxel: float = 0.5, nu: int = 384, pitch: float = 1.0):
    return {
        "SourceInit": [0, 1000.0, 0], "DetectorInit": [0, -500.0, 0], "StartAngle": 0,
        "EndAngle": 2 * np.pi, "NumberOfDetectorPixels": [nu, nu], "DetectorPixelSize": [pitch, pitch],
        "NumberOfViews": views, "ImagePixelSpacing": [voxel] * 3, "NumberOfImage": [n, n, n],
        "FilterType": "shepp-logan", "cutoff": 1, "DetectorShape": "Flat",
    }


class PairFactory:
    """Seeded (input, target, phantom) volumes for the sparse-view benchmark, on the GPU."""

    def __init__(self, device="cuda", n: int = 256, sparse_views: int = 60, full_views: int = 720,
                 sart_iters: int = 12, relaxation: float = 0.5, photons: float = 1e5,
                 n_ellipsoids: int = 20, backprojector: str = "matched"):
        self.device = torch.device(device)
        self.sparse = ScanGeometry(scan_params(sparse_views, n))
        self.full = ScanGeometry(scan_params(full_views, n))
        self.sart_iters = sart_iters
        self.relaxation = relaxation
        self.photons = photons
        self.n_ellipsoids = n_ellipsoids
        self.backprojector = backprojector

    def phantom(self, seed: int) -> torch.Tensor:
        ell = phantom.random_ellipsoids(self.n_ellipsoids, phantom.half_extent(self.sparse), seed)
        return phantom.ellipsoid_volume(self.sparse, ell, device=self.device) * phantom.MU_WATER

    def sparse_scan(self, vol: torch.Tensor, seed: int) -> torch.Tensor:
        gen = torch.Generator(device=self.device)
        gen.manual_seed(seed)
        return phantom.poisson_noise(P.dd_forward(vol, self.sparse), self.photons, generator=gen)

    def sart(self, proj: torch.Tensor, n_iter: int, x0=None) -> torch.Tensor:
        return IT.sart(proj, self.sparse, n_iter, relaxation=self.relaxation, x0=x0,
                       backprojector=self.backprojector)

    def pair(self, seed: int):
        vol = self.phantom(seed)
        p_sparse = self.sparse_scan(vol, seed + 1)
        x_in = self.sart(p_sparse, self.sart_iters)
        gen = torch.Generator(device=self.device)
        gen.manual_seed(seed + 2)
        p_full = phantom.poisson_noise(P.dd_forward(vol, self.full), self.photons, generator=gen)
        target = P.fdk(p_full, self.full)
        del p_full
        return x_in, target, vol


def train(net: PR.ResidualPrior, factory: PairFactory, n_pairs: int, steps_per_pair: int = 8,
          batch: int = 16, crop: int = 128, lr: float = 1e-3, seed: int = 0, log=None,
          opt: torch.optim.Optimizer | None = None, start_pair: int = 0, end_pair: int | None = None):
    """Train on pairs generated on the fly: pair k comes from phantom seed ``seed + 3 k``.

    The learning rate follows one cosine schedule over all ``n_pairs * steps_per_pair`` steps;
    pairs ``start_pair .. end_pair - 1`` are run, so that training can be split over several
    processes (pass the returned optimizer state back in).  Crop positions and slices of pair k
    come from a generator seeded with (seed, k), so a split run draws what one run would.
    Returns (optimizer, losses)."""
    if opt is None:
        opt = torch.optim.Adam(net.parameters(), lr=lr)
    end_pair = n_pairs if end_pair is None else min(end_pair, n_pairs)
    losses = []
    total = n_pairs * steps_per_pair
    for k in range(start_pair, end_pair):
        rng = np.random.default_rng([seed, k])
        x_in, target, _ = factory.pair(seed + 3 * k)
        nz, ny, nx = x_in.shape
        for j in range(steps_per_pair):
            step = k * steps_per_pair + j
            lr_now = lr * 0.5 * (1 + math.cos(math.pi * step / max(total, 1)))
            for gp in opt.param_groups:
                gp["lr"] = lr_now
            z = torch.from_numpy(rng.integers(1, nz - 1, size=batch)).to(x_in.device)
            y0 = int(rng.integers(0, ny - crop + 1))
            x0 = int(rng.integers(0, nx - crop + 1))
            inp = PR.slices_to_input(x_in, z)[:, y0:y0 + crop, x0:x0 + crop] * SCALE
            tgt = target[z][:, y0:y0 + crop, x0:x0 + crop] * SCALE
            out = net(inp.contiguous())
            loss = torch.mean((out - tgt) ** 2)
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            losses.append(float(loss.detach()))
        if log is not None:
            log(k, float(np.mean(losses[-steps_per_pair:])))
    return opt, losses


def residual(geom: ScanGeometry, x: torch.Tensor, proj: torch.Tensor) -> float:
    return float(torch.sqrt(torch.mean((P.dd_forward(x.contiguous(), geom) - proj) ** 2)))


def learned_reconstruction(factory: PairFactory, net: PR.ResidualPrior, proj: torch.Tensor,
                           dc_steps: int = 5, fallback_iters: int = 0):
    """SART warm start -> network -> ``dc_steps`` SART data-consistency steps.

    Returns (volume, info).  If the result fits ``proj`` worse than the warm start (projection
    residual), plain SART continues from the warm start for ``fallback_iters`` iterations and
    that is returned instead (info['fallback'] = True)."""
    g = factory.sparse
    x_warm = factory.sart(proj, factory.sart_iters)
    x_net = PR.apply_to_volume(net, x_warm, scale=SCALE)
    x = factory.sart(proj, dc_steps, x0=x_net)
    r_learned, r_warm = residual(g, x, proj), residual(g, x_warm, proj)
    info = {"residual_learned": r_learned, "residual_warm": r_warm, "fallback": False}
    if r_learned > r_warm:
        info["fallback"] = True
        x = factory.sart(proj, fallback_iters, x0=x_warm) if fallback_iters else x_warm
    return x, info


def time_to_rmse(factory: PairFactory, net: PR.ResidualPrior, seed: int, plain_iters: int = 200,
                 dc_steps: int = 5):
    """Wall time of plain SART (``plain_iters`` iterations) against the learned pipeline, and
    the RMSE (vs the phantom) of both.

    Plain SART runs twice: once with a per-iteration RMSE callback (untimed) to record its RMSE
    curve, once without callback for the wall time.  ``speedup`` compares with the full
    ``plain_iters`` run; ``speedup_to_rmse`` with the share of it up to the first iteration
    whose RMSE is already the final one (every iteration does the same work)."""
    vol = factory.phantom(seed)
    proj = factory.sparse_scan(vol, seed + 1)
    g = factory.sparse
    rmse = lambda x: float(torch.sqrt(torch.mean((x - vol) ** 2)))  # noqa: E731
    errs = []
    x_plain = IT.sart(proj, g, plain_iters, relaxation=factory.relaxation,
                      backprojector=factory.backprojector, callback=lambda it, xi: errs.append(rmse(xi)))
    target = errs[-1]
    it_target = next(i for i, e in enumerate(errs) if e <= target) + 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    IT.sart(proj, g, plain_iters, relaxation=factory.relaxation, backprojector=factory.backprojector)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    learned_reconstruction(factory, net, proj, dc_steps)  # warm-up: kernel loading is not timed
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x_l, info = learned_reconstruction(factory, net, proj, dc_steps)
    torch.cuda.synchronize()
    t_learned = time.perf_counter() - t0
    r_l = rmse(x_l)
    return {"plain_iters": plain_iters, "plain_rmse": target, "plain_time_s": t_plain,
            "plain_iters_to_rmse": it_target, "plain_time_to_rmse_s": t_plain * it_target / plain_iters,
            "learned_rmse": r_l, "learned_time_s": t_learned, "learned_reaches_plain_rmse": r_l <= target,
            "speedup": t_plain / t_learned,
            "speedup_to_rmse": t_plain * it_target / plain_iters / t_learned,
            "residual_plain": residual(g, x_plain, proj),
            "plain_rmse_curve": {i: errs[i - 1] for i in (1, 5, 10, 20, 25, 50, 100, 150, 200) if i <= len(errs)},
            **info}


def save_weights(net: PR.ResidualPrior, path: str, meta: dict | None = None) -> None:
    from safetensors.torch import save_file
    save_file({k: v.detach().float().cpu().contiguous() for k, v in net.state_dict().items()}, path,
              metadata={k: str(v) for k, v in (meta or {}).items()})


def weights_metadata(path: str) -> dict:
    """The training record stored with the weights (string values)."""
    from safetensors import safe_open
    with safe_open(path, framework="pt") as f:
        return dict(f.metadata() or {})


def pipeline_for(path: str, device="cuda"):
    """(factory, net, dc_steps) of the pipeline the weights at ``path`` were trained for: the
    warm start's SART iteration count and the data-consistency steps come from their metadata
    (20 and 5 when absent)."""
    meta = weights_metadata(path)
    factory = PairFactory(device=device, sart_iters=int(meta.get("sart_iters", 20)),
                          backprojector=meta.get("backprojector", "matched"))
    return factory, load_weights(path, device), int(meta.get("dc_steps", 5))


def load_weights(path: str, device="cuda") -> PR.ResidualPrior:
    from safetensors.torch import load_file
    net = PR.ResidualPrior()
    net.load_state_dict(load_file(path))
    return net.to(device)


def save_checkpoint(net: PR.ResidualPrior, opt: torch.optim.Optimizer, next_pair: int, path: str) -> None:
    torch.save({"net": {k: v.detach().cpu() for k, v in net.state_dict().items()},
                "opt": opt.state_dict(), "next_pair": next_pair}, path)


def load_checkpoint(path: str, lr: float, device="cuda"):
    ck = torch.load(path, map_location="cpu", weights_only=True)
    net = PR.ResidualPrior().to(device)
    net.load_state_dict(ck["net"])
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    opt.load_state_dict(ck["opt"])
    return net, opt, int(ck["next_pair"])


def main(argv=None):
    """Train the prior on pairs generated on the fly and evaluate time-to-RMSE on a held-out
    phantom.  ``--until`` and ``--resume`` split the training over several processes."""
    import argparse
    import json

    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--steps-per-pair", type=int, default=8)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default="prior.safetensors")
    ap.add_argument("--checkpoint", default=None, help="write a resumable checkpoint here")
    ap.add_argument("--resume", default=None, help="continue from this checkpoint")
    ap.add_argument("--until", type=int, default=None, help="stop before this pair index")
    ap.add_argument("--eval-seed", type=int, default=10000)
    ap.add_argument("--no-eval", action="store_true")
    ap.add_argument("--backprojector", default="matched", choices=["ref", "edge", "matched"])
    ap.add_argument("--sart-iters", type=int, default=12, help="SART iterations of the warm start")
    ap.add_argument("--dc-steps", type=int, default=5, help="SART data-consistency steps after the network")
    args = ap.parse_args(argv)
    factory = PairFactory(backprojector=args.backprojector, sart_iters=args.sart_iters)
    if args.resume:
        net, opt, start = load_checkpoint(args.resume, args.lr)
    else:
        net, opt, start = PR.ResidualPrior(seed=0).cuda(), None, 0
    t0 = time.perf_counter()
    log = lambda k, l: print(json.dumps({"pair": k, "loss": l, "t": time.perf_counter() - t0}), flush=True)  # noqa: E731
    opt, _ = train(net, factory, args.pairs, args.steps_per_pair, args.batch, args.crop, args.lr, seed=1,
                   log=log, opt=opt, start_pair=start, end_pair=args.until)
    t_train = time.perf_counter() - t0
    end = args.pairs if args.until is None else min(args.until, args.pairs)
    if args.checkpoint:
        save_checkpoint(net, opt, end, args.checkpoint)
    save_weights(net, args.out, {"pairs": end, "steps_per_pair": args.steps_per_pair,
                                 "batch": args.batch, "crop": args.crop, "lr": args.lr,
                                 "backprojector": args.backprojector, "sart_iters": args.sart_iters,
                                 "dc_steps": args.dc_steps})
    print(json.dumps({"trained_pairs": end, "train_s_this_run": t_train}), flush=True)
    if not args.no_eval:
        res = time_to_rmse(factory, net, args.eval_seed, dc_steps=args.dc_steps)
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
