"""Benchmark: FDK reconstruction of the full-size cone-beam scan on one MI355X.

Workload (default ``--config c3``): 512^3 volume (0.25 mm voxels) of 40 seeded random
ellipsoids (seed 3) times mu_water = 0.02/mm, 720 views over 360 degrees, 1024 x 768 flat
detector at 0.4 mm, SID 1000 mm, SDD 1500 mm, Shepp-Logan window; projections from the DD
forward projector plus Poisson noise at 1e5 photons.  All inputs are resident in HBM before
timing.  One step = one FDK reconstruction: cosine weight + ramp filter of every detector
row, then the FDK-weighted DD backprojection of all 720 views into the 512^3 volume.  The
forward projection is timed separately and reported beside the step.

Prints one JSON line.  ``value`` is seconds per FDK reconstruction (lower is better);
``roofline`` prices the dominant kernel (the backprojection) against the 8 TB/s HBM peak
with algorithmic bytes 4*(views*nv*nu + nx*ny*nz) per launch; ``cpu_baseline`` times the oracle's C
restatement (FP64, OpenMP on every host core) on 8 views spread over the scan into the whole
volume and scales the wall time to 720 views; a one-thread run of one view is reported beside it.
The forward projection's own CPU baseline (``dd_forward.cpu_baseline``) times the C restatement
of the oracle's slab loop on every cell of 4 views and scales it the same way.

``pixel_fdk`` times the reference's other FDK, the pixel-driven one (voxel-driven bilinear
backprojection, python/ConeBeamGPU.py:203-321), on the reference's geometry file
(python/ReconstructionParams.txt: 512^3, 680 views, 1024 x 768 detector at 0.390625 mm), with
its own CPU port (the pixel oracle's C restatement) timed beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the dominant kernel of one FDK step: the backprojection's common-case pass (the call also
# launches a flag fill and the remainder pass, both under 1% of it for this workload; the HIP
# events bracket all three, so the achieved rate is a lower bound for the pass itself)
BP_KERNEL = "dd_backproject_tile_kernel<32, 0>"
FP_KERNEL = "dd_forward_kernel"

# The reference's own geometry file for its pixel-driven FDK (python/ReconstructionParams.txt:1-11,
# read by python/ConeBeamGPU.py:164-197); AngleCoverage in radians as the reader stores it.
PIXEL_GEOMETRY = dict(SourceToAxis=1000.0, SourceToDetector=1500.0, DetectorPixelWidth=0.390625,
                      DetectorPixelHeight=0.390625, DetectorWidth=1024, DetectorHeight=768,
                      NumberOfViews=680, AngleCoverage=2 * np.pi, ReconX=512, ReconY=512, ReconZ=512,
                      DataPath="", precision="float32")
PIXEL_KERNEL = "pixel_backproject_kernel<32, 72>"

CONFIGS = {
    "c3": dict(n=512, voxel=0.25, views=720, nu=1024, nv=768, pitch=0.4, ellipsoids=40, seed=3,
               noise=True, workload="cone-beam FDK 512^3, 720 views, 1024x768 detector"),
    "c2": dict(n=256, voxel=0.5, views=360, nu=384, nv=384, pitch=1.0, ellipsoids=20, seed=2,
               noise=False, workload="cone-beam FDK 256^3, 360 views, 384x384 detector"),
    "c1": dict(n=256, nz=1, voxel=0.5, views=360, nu=512, nv=1, pitch=0.5, ellipsoids=10, seed=1,
               noise=False, shape="Curved", filter="ram-lak", ndim=2,
               workload="2D equiangular fan-beam FBP 256^2, 360 views, 512 bins"),
}


def make_params(cfg, filter_type="shepp-logan"):
    n = cfg["n"]
    return {
        "SourceInit": [0, 1000.0, 0], "DetectorInit": [0, -500.0, 0], "StartAngle": 0,
        "EndAngle": 2 * np.pi, "NumberOfDetectorPixels": [cfg["nu"], cfg["nv"]],
        "DetectorPixelSize": [cfg["pitch"], cfg["pitch"]], "NumberOfViews": cfg["views"],
        "ImagePixelSpacing": [cfg["voxel"]] * 3, "NumberOfImage": [n, n, cfg.get("nz", n)],
        "PhantomCenter": [0, 0, 0], "RotationOrigin": [0, 0, 0], "ReconCenter": [0, 0, 0],
        "Method": "Distance", "FilterType": cfg.get("filter", filter_type), "cutoff": 1, "GPU": 1,
        "DetectorShape": cfg.get("shape", "Flat"), "Pitch": 0, "DetectorOffset": [0, 0],
    }


def profile_record(name, kernel):
    """Counter record of rocprofv3 --pmc runs, profiles/<name>.json: HBM bytes per launch
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) and texture-unit busy
    fractions; {} when absent or when the record is for another kernel."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        rec = json.load(f)
    return rec if rec.get("kernel") == kernel else {}


def cpu_baseline(params, raw, geom, sample=8, check=4):
    """Time the CPU port of the FDK step and scale it to a full reconstruction.

    The port is the oracle's C restatement (``oracle/c/dd_backproject.c``: FP64, the oracle's
    operations in its order, OpenMP threads over the image rows) with the oracle's filter batched
    over views on ``scipy.fft`` workers, both on every host core this process may use.  Sample:
    filter + backprojection of ``sample`` views spread evenly over the scan into the whole volume;
    the wall time is scaled by nview / sample (each view of a circular scan is the same work).
    ``check`` views are timed as well to show that the per-view time holds, and one view on one
    thread gives the single-core figure.  The sample's volume and filtered rows are compared with
    the GPU's filter and backprojection of the same views."""
    import ctreconstruction_amd as ctr
    from oracle import dd_c as C
    from oracle import dd_oracle as O

    g = O.Geometry(params)
    threads = C.host_threads()
    C.lib()

    def spread(k):
        return [int(round(j * g.nview / k)) for j in range(k)]

    def timed(views, t):
        r = raw[views].cpu().numpy()
        t0 = time.perf_counter()
        q = C.filter_projections(g, r, threads=t)
        t1 = time.perf_counter()
        vol = C.backproject(g, q, "ref", views=views, threads=t)
        t2 = time.perf_counter()
        return q, vol, t1 - t0, t2 - t1

    views = spread(sample)
    timed(views[:1], threads)  # untimed: FFT plans and the thread pool start here
    q_c, vol_c, t_f, t_bp = timed(views, threads)
    full = (t_f + t_bp) * g.nview / sample
    _, _, t_f_chk, t_bp_chk = timed(spread(check), threads)
    full_chk = (t_f_chk + t_bp_chk) * g.nview / check
    _, _, t_f_1, t_bp_1 = timed(views[:1], 1)
    single = (t_f_1 + t_bp_1) * g.nview
    q_s = ctr.fdk_filter(raw[views].contiguous(), geom)
    vol_gpu = ctr.dd_backproject(q_s, geom.with_views(views)).cpu().numpy()
    err = float(np.abs(vol_gpu - vol_c).max() / np.abs(vol_c).max())
    err_filter = float(np.abs(q_s.cpu().numpy() - q_c).max() / np.abs(q_c).max())
    return {
        "value": full, "unit": "s", "cores": threads, "kind": "port",
        "sample": (f"C restatement of the FP64 oracle (gcc -O3, OpenMP) on {threads} threads: filter "
                   f"and FDK backprojection of {sample} views spread over the scan into the whole "
                   f"{g.nx}x{g.ny}x{g.nz} volume in {t_f + t_bp:.2f} s wall (filter {t_f:.2f} s, "
                   f"backprojection {t_bp:.2f} s), scaled linearly to {g.nview} views"),
        "cpu_seconds_in_sample": (t_f + t_bp) * threads,
        "extrapolation_check": {f"{check}_views_s": full_chk, f"{sample}_views_s": full,
                                "ratio": full / full_chk},
        "single_thread_s": single,
        "single_thread_sample": (f"1 view filtered and backprojected on 1 thread in {t_f_1 + t_bp_1:.2f} s "
                                 f"(filter {t_f_1:.3f} s), scaled to {g.nview} views"),
        "host_cpus": os.cpu_count(), "host_affinity": threads,
        "sample_error_vs_oracle": {"fdk_backprojection": err, "filter": err_filter},
    }


def forward_cpu_baseline(params, vol, gpu_views, views):
    """CPU port of the forward projection: the oracle's C restatement of its slab loop
    (``oracle/c/dd_forward.c``, FP64, OpenMP over the detector cells on every host core) on
    ``views`` spread over the scan, scaled by nview / len(views); its projections are compared with
    the GPU's (noise-free) projections of the same views."""
    from oracle import dd_c as C
    from oracle import dd_oracle as O

    g = O.Geometry(params)
    threads = C.host_threads()
    C.lib()
    v = vol.cpu().numpy()
    t0 = time.perf_counter()
    ref = C.forward_project(g, v, views=views, threads=threads)
    t = time.perf_counter() - t0
    return {"value": t * g.nview / len(views), "unit": "s", "cores": threads, "kind": "port",
            "sample": (f"C restatement of the FP64 oracle's forward slab loop (gcc -O3, OpenMP) on {threads} "
                       f"threads: every cell of {len(views)} views spread over the scan in {t:.2f} s wall, "
                       f"scaled to {g.nview} views"),
            "sample_error_vs_oracle": float(np.abs(gpu_views - ref).max() / np.abs(ref).max())}


WEIGHTS = os.path.join(ROOT, "ctreconstruction_amd", "weights", "prior_sparse60.safetensors")


def pixel_fdk_leg(dev, stream, steps, warmup, cpu=True, seed=3):
    """Pixel-driven FDK (``ConeBeam.Reconstruction``, python/ConeBeamGPU.py:203-321) on the
    reference's geometry file: weight + Ram-Lak (cutoff 0.3) filter of every row, then the
    voxel-driven bilinear backprojection of 680 views into 512^3, event-timed on the stream.
    Projections: the seeded random-ellipsoid phantom, DD-forward-projected on the same geometry.
    ``cpu``: the C port of the pixel oracle on every host core (:func:`pixel_cpu_baseline`)."""
    import torch

    import ctreconstruction_amd as ctr
    from ctreconstruction_amd import paramsfile, phantom
    from ctreconstruction_amd import pixel_fdk as PF

    scan = PF.PixelScan(PIXEL_GEOMETRY)
    geom = ctr.ScanGeometry(paramsfile.to_recon_params(PIXEL_GEOMETRY))
    ell = phantom.random_ellipsoids(40, phantom.half_extent(geom), seed)
    vol = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    proj = ctr.dd_forward(vol, geom)
    del vol
    spec = PF.pixel_spectrum(scan.zero_padded_length, scan.tau, device=dev)
    q = torch.empty_like(proj)
    rec = torch.empty(scan.vol_shape, dtype=torch.float32, device=dev)
    ang = scan.device_angles(dev)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        PF.pixel_filter(proj, scan, spec, out=q)
        if ev is not None:
            ev[1].record(stream)
        PF.pixel_backproject(q, scan, out=rec, angles=ang)
        if ev is not None:
            ev[2].record(stream)

    for _ in range(warmup):
        step()
    evs = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(steps)]
    for k in range(steps):
        step(evs[k])
    torch.cuda.synchronize()
    f_s = float(np.mean([a.elapsed_time(b) for a, b, _ in evs])) / 1e3
    b_s = float(np.mean([b.elapsed_time(c) for _, b, c in evs])) / 1e3
    updates = float(scan.nx * scan.ny * scan.nz) * scan.n_views
    prof = profile_record("pixel_backproject_traffic", PIXEL_KERNEL)
    out = {"workload": "pixel-driven FDK 512^3, 680 views, 1024x768 detector (python/ReconstructionParams.txt)",
           "s": f_s + b_s, "filter_s": f_s, "backprojection_s": b_s, "gups": updates / b_s / 1e9,
           "kernel": PIXEL_KERNEL, "bound": "valu issue",
           "valu_issue_per_simd_cycle": prof.get("valu_issue_per_simd_cycle"),
           "hbm_traffic_bytes": prof.get("traffic_bytes")}
    if cpu:
        out["cpu_baseline"] = pixel_cpu_baseline(proj, q, scan, ang)
    return out


def pixel_cpu_baseline(proj, q, scan, ang, sample=8, check=4):
    """CPU port of the pixel-driven FDK, timed as :func:`cpu_baseline` times the DD one: the
    pixel oracle's C restatement (``oracle/c/pixel_backproject.c``, FP64, OpenMP) and its filter
    batched on ``scipy.fft``, on every host core; ``sample`` views spread over the scan into the
    whole volume, scaled by nview / sample; ``check`` views and one view on one thread beside it.
    The GPU's backprojection of the same views (of the GPU's filtered rows) is compared with the
    port's on the voxels whose sample lies more than 2e-3 cells from a bounds-test flip."""
    from ctreconstruction_amd import pixel_fdk as PF
    from oracle import dd_c as C
    from oracle import pixel_c as PC
    from oracle import pixel_oracle as PO

    og = PO.PixelGeometry(PIXEL_GEOMETRY)
    threads = C.host_threads()
    C.lib()

    def spread(k):
        return [int(round(j * og.ns / k)) for j in range(k)]

    def timed(views, t):
        r = proj[views].cpu().numpy()
        t0 = time.perf_counter()
        qs = PC.filter_projections(og, r, threads=t)
        t1 = time.perf_counter()
        vol = PC.backproject(og, qs, views, threads=t)
        return qs, vol, t1 - t0, time.perf_counter() - t1

    views = spread(sample)
    timed(views[:1], threads)  # untimed: FFT plans and the thread pool start here
    q_c, _, t_f, t_bp = timed(views, threads)
    full = (t_f + t_bp) * og.ns / sample
    _, _, t_f_chk, t_bp_chk = timed(spread(check), threads)
    full_chk = (t_f_chk + t_bp_chk) * og.ns / check
    _, _, t_f_1, t_bp_1 = timed(views[:1], 1)
    q_g = q[views].cpu().numpy()
    vol_c, margin = PC.backproject(og, q_g, views, threads=threads, margin=True)
    sub = PF.PixelScan({**PIXEL_GEOMETRY, "NumberOfViews": sample})
    vol_g = PF.pixel_backproject(q[views].contiguous(), sub, scale=scan.dtheta,
                                 angles=ang[views].contiguous()).cpu().numpy()
    keep = margin > 2e-3
    return {
        "value": full, "unit": "s", "cores": threads, "kind": "port",
        "sample": (f"C restatement of the FP64 pixel oracle (gcc -O3, OpenMP) on {threads} threads: "
                   f"filter and backprojection of {sample} views spread over the scan into the whole "
                   f"512^3 volume in {t_f + t_bp:.2f} s wall (filter {t_f:.2f} s), scaled to {og.ns} views"),
        "extrapolation_check": {f"{check}_views_s": full_chk, f"{sample}_views_s": full, "ratio": full / full_chk},
        "single_thread_s": (t_f_1 + t_bp_1) * og.ns,
        "sample_error_vs_oracle": {
            "backprojection": float(np.abs(vol_g - vol_c)[keep].max() / np.abs(vol_c).max()),
            "filter": float(np.abs(q_g - q_c).max() / np.abs(q_c).max()),
            "voxels_compared": float(keep.mean())},
    }


CALIB_KINDS = ("v_fma_f32", "v_add_u32", "v_cndmask_b32", "v_readfirstlane_b32", "v_rcp_f32", "v_add_f64",
               "v_mul_u32_u24")
SIMDS = 1024          # 256 CUs x 4 SIMDs
NOMINAL_CLOCK_HZ = 2.4e9


def valu_calibration(dev, stream, blocks=4096, iters=4096, reps=3):
    """Issue-rate microbenchmarks (ctreconstruction_amd/csrc/calibrate.hip): for each instruction
    kind, kernels whose wavefronts issue exactly 8 * iters instructions of that kind, event-timed.
    The rate is wave-instructions per SIMD-cycle at the nominal 2.4 GHz clock; a rocprofv3 pass over the same run reads SQ_INSTS_VALU and
    GRBM_GUI_ACTIVE for the measured clock."""
    import ctypes

    import torch

    from ctreconstruction_amd import _native as N

    lib = ctypes.CDLL(N.CALIB_PATH)
    lib.ctr_calib_run.restype = ctypes.c_int
    lib.ctr_calib_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    sink = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    st = ctypes.c_void_p(stream.cuda_stream)
    out = {"blocks": blocks, "threads_per_block": 256, "iters": iters, "clock_hz_assumed": NOMINAL_CLOCK_HZ}
    waves = blocks * 4
    for kind, name in enumerate(CALIB_KINDS):
        per_iter = 8
        if lib.ctr_calib_run(kind, blocks, iters, ctypes.c_void_p(sink.data_ptr()), st) != 0:
            raise RuntimeError(f"calibration kernel {name} failed to launch")
        ts = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            lib.ctr_calib_run(kind, blocks, iters, ctypes.c_void_p(sink.data_ptr()), st)
            b.record(stream)
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b) / 1e3)
        t = float(np.median(ts))
        wave_instr = float(waves) * iters * per_iter
        out[name] = {"s": t, "wave_instructions": wave_instr,
                     "per_simd_cycle": wave_instr / (SIMDS * NOMINAL_CLOCK_HZ * t)}
    return out


HELD_OUT_SEEDS = (10000, 10003, 10006)  # training used phantom seeds 1 .. 6000


def learned_vs_sart(seeds=HELD_OUT_SEEDS):
    """Sparse-view benchmark (256^3, 60 views, 384^2, Poisson 1e5, held-out phantoms): wall time
    and RMSE vs the phantom of plain 200-iteration SART and of the learned prior + SART pipeline
    (trained weights in ctreconstruction_amd/weights/), per phantom and summarised (min and mean
    of both ratios); None when no trained weights exist."""
    if not os.path.exists(WEIGHTS):
        return None
    from ctreconstruction_amd import learned as L
    factory, net, dc_steps = L.pipeline_for(WEIGHTS)
    per_seed = []
    for seed in seeds:
        res = L.time_to_rmse(factory, net, seed, dc_steps=dc_steps)
        res["phantom_seed"] = seed
        per_seed.append(res)
    sp = [r["speedup"] for r in per_seed]
    spr = [r["speedup_to_rmse"] for r in per_seed]
    return {"phantom_seeds": list(seeds),
            "pipeline": f"SART-{factory.sart_iters} warm start, network, {dc_steps} SART data-consistency steps",
            "speedup_min": min(sp), "speedup_mean": float(np.mean(sp)),
            "speedup_to_rmse_min": min(spr), "speedup_to_rmse_mean": float(np.mean(spr)),
            "learned_reaches_plain_rmse_all": all(r["learned_reaches_plain_rmse"] for r in per_seed),
            "per_phantom": per_seed}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-prior time-to-RMSE leg")
    ap.add_argument("--no-pixel", action="store_true", help="skip the pixel-driven FDK leg")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (rocprof counter calibration)")
    ap.add_argument("--calibrate-valu", action="store_true",
                    help="also run the instruction-issue microbenchmarks (VALU ceiling calibration)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; the view-sharded multi-GPU layout is "
                                   "designed, not built", "n_gpus": args.gpus}))
        return 2
    import torch

    import ctreconstruction_amd as ctr
    from ctreconstruction_amd import phantom

    cfg = CONFIGS[args.config]
    params = make_params(cfg)
    geom = ctr.ScanGeometry(params)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)

    t_setup = time.perf_counter()
    ell = phantom.random_ellipsoids(cfg["ellipsoids"], phantom.half_extent(geom), cfg["seed"],
                                    ndim=cfg.get("ndim", 3))
    vol = phantom.ellipsoid_volume(geom, ell, device=dev) * phantom.MU_WATER
    # forward projection (timed separately from the FDK step)
    proj = ctr.dd_forward(vol, geom)
    torch.cuda.synchronize()
    fp_times = []
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctr.dd_forward(vol, geom, out=proj)
        e1.record(stream)
        torch.cuda.synchronize()
        fp_times.append(e0.elapsed_time(e1) / 1e3)
    fp_views = [int(round(j * geom.nview / 4)) for j in range(4)]
    fp_sample = proj[fp_views].cpu().numpy()  # noise-free, for the forward CPU baseline's check
    if cfg["noise"]:
        gen = torch.Generator(device=dev)
        gen.manual_seed(cfg["seed"])
        proj = phantom.poisson_noise(proj, 1e5, generator=gen)
    spectrum = ctr.filter_spectrum(geom, device=dev)
    q = torch.empty_like(proj)
    rec = torch.empty(geom.vol_shape, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    def step(ev=None):
        ctr.fdk_filter(proj, geom, spectrum, out=q)
        if ev is not None:
            ev[0].record(stream)
        ctr.dd_backproject(q, geom, mode="ref", out=rec)
        if ev is not None:
            ev[1].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    bp_events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                 for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(bp_events[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    bp_s = float(np.mean([a.elapsed_time(b) for a, b in bp_events])) / 1e3
    sec = elapsed / args.steps

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    calib = valu_calibration(dev, stream) if args.calibrate_valu else None

    nvox = geom.nx * geom.ny * geom.nz
    ncell = geom.nview * geom.nv * geom.nu
    bp_bytes = 4.0 * (ncell + nvox)
    achieved = bp_bytes / bp_s / 1e9
    updates = float(nvox) * geom.nview
    fp_s = float(np.median(fp_times))
    ray_slab = float(ncell) * geom.ny  # every ray crosses n_slab slabs (ny == nx here)
    bp_prof = profile_record(f"dd_backproject_{args.config}_traffic", BP_KERNEL)
    fp_prof = profile_record(f"dd_forward_{args.config}_traffic", FP_KERNEL)
    out = {
        "metric": "FDK 512³/720-view recon s; DD proj GUPS; SART vs learned time-to-RMSE",
        "value": sec, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded random ellipsoids, Poisson noise 1e5 photons)"
        if cfg["noise"] else "synthetic (seeded random ellipsoids)",
        "config": {"workload": cfg["workload"], "volume": [geom.nz, geom.ny, geom.nx],
                   "voxel_mm": cfg["voxel"], "views": geom.nview, "detector": [geom.nv, geom.nu],
                   "pixel_mm": cfg["pitch"], "sad_mm": geom.sad, "sdd_mm": geom.sdd,
                   "filter": geom.filter_type, "step": "fdk_filter + dd_backproject (mode ref)"},
        "roofline": {"bound": "hbm", "kernel": BP_KERNEL, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "algorithmic_bytes": bp_bytes, "traffic": bp_prof.get("traffic_bytes"),
                     "texture_busy": {"ta": bp_prof.get("ta_busy_frac"), "td": bp_prof.get("td_busy_frac")},
                     # what bounds it: instruction issue (the measured lone-instruction ceilings of
                     # --calibrate-valu: v_fma_f32 0.28, v_add_u32 0.43 wave-instructions per
                     # SIMD-cycle), with the LDS busy beside it; not HBM
                     "limiter": "instruction issue (VALU), LDS beside it",
                     "valu_issue_per_simd_cycle": bp_prof.get("valu_issue_per_simd_cycle"),
                     "valu_lane_instructions_per_update": bp_prof.get("valu_lane_instructions_per_update"),
                     "lds_active_frac": bp_prof.get("lds_active_frac")},
        "fdk_backprojection": {"s": bp_s, "gups": updates / bp_s / 1e9},
        "fdk_filter": {"s": sec - bp_s},
        "dd_forward": {"s": fp_s, "gups": updates / fp_s / 1e9, "ray_slab_updates_per_s": ray_slab / fp_s,
                       "hbm_traffic_bytes": fp_prof.get("traffic_bytes"),
                       "texture_busy": {"ta": fp_prof.get("ta_busy_frac"), "td": fp_prof.get("td_busy_frac")},
                       "valu_issue_per_simd_cycle": fp_prof.get("valu_issue_per_simd_cycle"),
                       "limiter": "LDS (staged z segments and their breakpoint reads)",
                       "lds_active_frac": fp_prof.get("lds_active_frac")},
        "sart_vs_learned_time_to_rmse": None if args.no_learned else learned_vs_sart(),
        "pixel_fdk": None if args.no_pixel else pixel_fdk_leg(dev, stream, args.steps, args.warmup,
                                                             cpu=not args.no_cpu_baseline),
        "setup_s": t_setup,
    }
    if calib is not None:
        out["valu_calibration"] = calib
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(params, proj, geom)
        out["dd_forward"]["cpu_baseline"] = forward_cpu_baseline(params, vol, fp_sample, fp_views)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
