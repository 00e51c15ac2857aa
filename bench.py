#!/usr/bin/env python3
"""Benchmark: cone-beam FDK on a 384^3 volume with 512 views on one MI355X.

One step = CUDAmex_FP (cone, step 1.0, clamp) -> FDK pre-filter (cosine
weight, Hamming-windowed ramp, angle step / 2) -> CUDAmex_BP (cone), plain
FP32, inputs resident in HBM.  The headline value is FP+BP GUPS:
  (U_FP + U_BP) / (t_FP + t_BP) / 1e9,  U = 384^3 * 512 = 2.899e10 per
operator (every voxel, every view).  The same line reports each operator's
GUPS, the FP16-storage form's GUPS and its relative RMSE against the FP64
oracle on a sample of rays/voxel columns, the roofline of the FP kernel
and a CPU baseline (the oracle on a bounded sample, this host's cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mex-cuda-ct_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))  # the checker (cpu_baseline, FP64 sample)

HBM_PEAK_GBS = 8000.0
VC_PEAK_GBS = 256 * 64 * 2.4  # vector L1 caches: 64 B/clock per CU, 256 CUs, 2.4 GHz
# Memory-side bytes of one cone 384^3 x 512 f32 FP launch from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE (copy-calibrated): what the L2s fetched from
# and wrote to beyond themselves (the memory-side cache or HBM).  Valid for
# the kernel sources whose hash is stamped beside it; a changed source
# reports null until the counters are taken again.
FP_TRAFFIC = {"bytes": 56.71e9, "sources_sha16": "6750af7b65171f42"}
FP_KERNEL_SOURCES = ("mex-cuda-ct_amd/csrc/fpz.hip", "mex-cuda-ct_amd/csrc/ray.h",
                     "mex-cuda-ct_amd/csrc/common.h")


def kernel_sources_sha16():
    import hashlib
    h = hashlib.sha256()
    for f in FP_KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def host_cpu():
    """(cores this process may run on, the host's logical CPUs, CPU model)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return avail, os.cpu_count() or avail, model


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=384)
    ap.add_argument("--views", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-f16", action="store_true", help="skip the FP16-storage form")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the parallel-beam, general-3D and OSC-TV workload timings")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 1 GiB device-to-device copy (the reference kernel for "
                         "scaling HBM byte counters)")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=10.0, help="Poisson arrivals per second")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)

    import torch

    import mexct
    from mexct import phantoms
    from mexct.device import Context

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    W = H = args.size
    n = args.views
    SAD = 1000.0
    log(f"[bench] building inputs W=H={W} views={n}")
    vol_np = phantoms.cylinder(W, H, 140.0 * W / 384.0, 0.02)
    ang = phantoms.angles_linspace(n, 2 * np.pi)
    vol = torch.from_numpy(vol_np.T.copy()).to(dev)  # (H, W, W) == MATLAB [W, W, H]
    sino = torch.empty((n, H, W), dtype=torch.float32, device=dev)
    filt = torch.empty_like(sino)
    rec = torch.empty((H, W, W), dtype=torch.float32, device=dev)
    # the operators run on `stream`; the events are recorded on it too
    stream = torch.cuda.Stream(device=dev)
    ctx = Context(0, stream=stream, order_with_torch=False)
    torch.cuda.synchronize()
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)  # reads and writes exactly 1 GiB
        torch.cuda.synchronize()
        del src, dst

    def run_step(storage, v, events=None):
        if events is not None:
            events[0].record(stream)
        ctx.fp(v, SAD, ang, n, 2, sino, storage=storage)
        if events is not None:
            events[1].record(stream)
        ctx.prefilter(sino, 2, SAD, ang, 3, 1.0, filt)
        if events is not None:
            events[2].record(stream)
        ctx.bp(filt, SAD, ang, 2, rec, storage=storage)
        if events is not None:
            events[3].record(stream)

    def timed(storage, v, steps, warmup):
        for _ in range(warmup):
            run_step(storage, v)
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
        t0 = time.perf_counter()
        for k in range(steps):
            run_step(storage, v, evs[k])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        fp_ms = np.mean([e[0].elapsed_time(e[1]) for e in evs])
        fl_ms = np.mean([e[1].elapsed_time(e[2]) for e in evs])
        bp_ms = np.mean([e[2].elapsed_time(e[3]) for e in evs])
        return (t1 - t0) * 1e3 / steps, fp_ms, fl_ms, bp_ms

    U = float(W) * W * H * n
    log(f"[bench] fp32: warmup {args.warmup}, steps {args.steps}")
    step_ms, fp_ms, fl_ms, bp_ms = timed(mexct.F32, vol, args.steps, args.warmup)
    fp_gups = U / (fp_ms * 1e-3) / 1e9
    bp_gups = U / (bp_ms * 1e-3) / 1e9
    value = 2 * U / ((fp_ms + bp_ms) * 1e-3) / 1e9
    log(f"[bench] fp32 step {step_ms:.2f} ms: FP {fp_ms:.2f} ms ({fp_gups:.0f} GUPS), "
        f"filter {fl_ms:.2f} ms, BP {bp_ms:.2f} ms ({bp_gups:.0f} GUPS)")
    sino32 = sino.clone()
    filt32 = filt.clone()
    rec32 = rec.clone()

    # Measured gather roofline: each operator against its own loads-only
    # form (the same launches, block order, addresses and loads, with the
    # taps XOR-folded instead of interpolated), which it cannot outrun.
    # Rates are texels gathered per second: 8 per FP sample, 4 per BP
    # update, storage bytes each; the vector caches' peak (64 B per clock
    # per CU, 256 CUs at 2.4 GHz) is reported beside.
    import oracle as O  # checker: sample-count estimate for the FP gather rate

    scratch = {"fp": torch.empty_like(sino), "bp": torch.empty_like(rec)}

    def loads_ms(kind, storage, v, reps=3):
        """Mean time of the loads-only FP (on volume v) or BP (on the filtered
        sinogram), measured like the operators: events around the call."""
        out = scratch[kind]
        if kind == "fp":
            call = lambda: ctx.fp_loads(v, SAD, ang, n, 2, out, storage=storage)  # noqa: E731
        else:
            call = lambda: ctx.bp_loads(filt32, SAD, ang, 2, out, storage=storage)  # noqa: E731
        call()  # warm
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            call()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    views8 = np.linspace(0, n - 1, 8).astype(int)
    S_est = float(np.mean([O.fp(vol_np, SAD, ang[v:v + 1], 1, 2, return_samples=True)[1]
                           for v in views8])) * n
    hw = W // 2 - 1
    gx, gy = np.meshgrid(np.arange(W), np.arange(W), indexing="ij")
    U_bp = float(((gx - hw) ** 2 + (gy - hw) ** 2 < hw * hw).sum()) * H * n

    def gather_fracs(storage, t_fp, t_bp, v):
        es = 4 if storage == mexct.F32 else 2
        lf = loads_ms("fp", storage, v)
        lb = loads_ms("bp", storage, v)
        out = {}
        for k, t, tl, nbytes in (("fp", t_fp, lf, S_est * 8 * es), ("bp", t_bp, lb, U_bp * 4 * es)):
            a = nbytes / (t * 1e-3) / 1e9
            p = nbytes / (tl * 1e-3) / 1e9
            out[k] = {"achieved_GBs": round(a, 1), "measured_peak_GBs": round(p, 1),
                      "frac": round(a / p, 4), "loads_only_ms": round(tl, 3),
                      "frac_of_vector_cache_peak": round(a / VC_PEAK_GBS, 4)}
        return out

    gather = gather_fracs(mexct.F32, fp_ms, bp_ms, vol)
    gather.update({"fp_samples": S_est,
                   "fp_samples_note": "ray-march samples, mean of 8 views x views (oracle count)",
                   "bp_updates_in_circle": U_bp,
                   "unit": "GB/s of gathered texels (8 per FP sample, 4 per BP update, "
                           "storage bytes each)",
                   "ceiling": "measured_peak = the operator's loads-only form (same launch, "
                              "every load, taps XOR-folded instead of interpolated)",
                   "vector_cache_peak_GBs": VC_PEAK_GBS})
    log(f"[bench] gather roofline: FP {gather['fp']['achieved_GBs']:.0f}/"
        f"{gather['fp']['measured_peak_GBs']:.0f} GB/s ({gather['fp']['frac']:.2f}), "
        f"BP {gather['bp']['achieved_GBs']:.0f}/{gather['bp']['measured_peak_GBs']:.0f} GB/s "
        f"({gather['bp']['frac']:.2f})")

    # FP roofline: compulsory HBM bytes of a view-by-view projector, the
    # volume read once per view plus the projections written once.
    fp_bytes = n * (float(W) * W * H * 4) + n * float(W) * H * 4
    fp_gbs = fp_bytes / (fp_ms * 1e-3) / 1e9
    # traffic: memory-side bytes per FP launch from rocprofv3 --pmc FETCH_SIZE
    # and WRITE_SIZE on this workload, scaled by what they read for a 1 GiB
    # device copy (FETCH x2.0, WRITE x1.0); rerun the counters when the kernel
    # changes.  It is below the algorithmic bytes: a row group's volume slab
    # stays in the memory-side cache across the views.
    sha = kernel_sources_sha16()
    fresh = sha == FP_TRAFFIC["sources_sha16"] and (W, n) == (384, 512)
    roofline = {"bound": "hbm", "kernel": "fpz_kernel<cone,f32,x-pair volume>",
                "achieved": round(fp_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(fp_gbs / HBM_PEAK_GBS, 4),
                "traffic": FP_TRAFFIC["bytes"] if fresh else None,
                "traffic_note": ("rocprofv3 FETCH_SIZE+WRITE_SIZE per launch, copy-calibrated, "
                                 f"taken on kernel sources {FP_TRAFFIC['sources_sha16']}"
                                 if fresh else
                                 f"kernel sources {sha} differ from the counter run's "
                                 f"{FP_TRAFFIC['sources_sha16']} (or another workload): "
                                 "no measured traffic"),
                "algorithmic_bytes_per_launch": fp_bytes,
                "bound_note": ("the kernel's own limit is the vector memory path (texture "
                               "addresser busy ~98% of the launch, rocprofv3 TA_TA_BUSY); "
                               "see gather_roofline.fp.frac_of_vector_cache_peak")}

    f16 = None
    if not args.no_f16:

        vol16 = vol.half()
        torch.cuda.synchronize()
        log("[bench] fp16-storage")
        s16, f16_fp_ms, _, f16_bp_ms = timed(mexct.F16, vol16, args.steps, 1)
        rng = np.random.Generator(np.random.PCG64(7))
        nr = 2048
        ix = rng.integers(0, W, nr)
        iy = rng.integers(0, H, nr)
        iv = rng.integers(0, n, nr)
        ref = O.fp_f64_rays(vol_np, SAD, ang, 2, ix, iy, iv)
        g16 = sino.cpu().numpy()[iv, iy, ix]
        g32 = sino32.cpu().numpy()[iv, iy, ix]
        rr16 = float(np.linalg.norm(g16 - ref) / np.linalg.norm(ref))
        rr32 = float(np.linalg.norm(g32 - ref) / np.linalg.norm(ref))
        # BP accuracy: same filtered FP32 input, columns inside the circle
        hw = W // 2 - 1
        cols = []
        while len(cols) < 48:
            a, b = rng.integers(0, W, 2)
            if (a - hw) ** 2 + (b - hw) ** 2 < hw * hw:
                cols.append((a, b))
        cx = np.array([c[0] for c in cols])
        cy = np.array([c[1] for c in cols])
        fil_np = np.asfortranarray(filt32.cpu().numpy().transpose(2, 1, 0))
        refb = O.bp_f64_cols(fil_np, SAD, ang, 2, cx, cy)  # [ncol, H]
        rec16 = rec.cpu().numpy()[:, cy, cx].T
        rec32n = rec32.cpu().numpy()[:, cy, cx].T
        bb16 = float(np.linalg.norm(rec16 - refb) / np.linalg.norm(refb))
        bb32 = float(np.linalg.norm(rec32n - refb) / np.linalg.norm(refb))
        gr16 = gather_fracs(mexct.F16, f16_fp_ms, f16_bp_ms, vol16)
        log(f"[bench] fp16 gather roofline: FP {gr16['fp']['frac']:.2f}, BP {gr16['bp']['frac']:.2f}")
        f16 = {"fp_ms": round(f16_fp_ms, 3), "bp_ms": round(f16_bp_ms, 3), "gather_roofline": gr16,
               "fp_gups": round(U / (f16_fp_ms * 1e-3) / 1e9, 1),
               "bp_gups": round(U / (f16_bp_ms * 1e-3) / 1e9, 1),
               "fp_speedup_vs_f32": round(fp_ms / f16_fp_ms, 3),
               "bp_speedup_vs_f32": round(bp_ms / f16_bp_ms, 3),
               "fp_rel_rmse_vs_f64": rr16, "bp_rel_rmse_vs_f64": bb16,
               "f32_fp_rel_rmse_vs_f64": rr32, "f32_bp_rel_rmse_vs_f64": bb32,
               "sample": f"{nr} random rays, {len(cols)} voxel columns"}
        log(f"[bench] fp16: FP {f16_fp_ms:.2f} ms BP {f16_bp_ms:.2f} ms; relRMSE FP {rr16:.2e} "
            f"(f32 {rr32:.2e}) BP {bb16:.2e} (f32 {bb32:.2e})")

    cpu = None
    if not args.no_cpu:
        import oracle as O

        avail, logical, model = host_cpu()
        # at most 16 threads: the host share of one GPU on the box
        cores = max(1, min(avail, 16))
        O.set_threads(cores)
        kv = 64  # views of the FP sample (about 5 s on 16 threads)
        log(f"[bench] cpu baseline: oracle FP on {kv} views, BP on 1024 columns, {cores} threads")
        t0 = time.perf_counter()
        O.fp(vol_np, SAD, ang[:kv], kv, 2)
        tf = time.perf_counter() - t0
        rng = np.random.Generator(np.random.PCG64(11))
        ncol = 32768  # voxel columns of the BP sample (about 3 s)
        cx = rng.integers(0, W, ncol)
        cy = rng.integers(0, W, ncol)
        fil_np = np.asfortranarray(filt32.cpu().numpy().transpose(2, 1, 0))
        t0 = time.perf_counter()
        O.bp_cols(fil_np, SAD, ang, 2, cx, cy)
        tb = time.perf_counter() - t0
        cfp = float(W) * W * H * kv / tf / 1e9
        cbp = float(ncol) * H * n / tb / 1e9
        cpu = {"value": round(2.0 / (1.0 / cfp + 1.0 / cbp), 4), "unit": "GUPS", "cores": cores,
               "kind": "port", "fp_gups": round(cfp, 4), "bp_gups": round(cbp, 4),
               "host_cpus_available": avail, "host_cpus_logical": logical, "cpu_model": model,
               "threads_note": "capped at 16 threads, the host share of one GPU",
               "sample": f"FP of {kv} of {n} views ({tf:.1f} s); BP of {ncol} random voxel "
                         f"columns x {H} z x {n} views ({tb:.1f} s); FP32 oracle"}
        log(f"[bench] cpu: FP {cfp:.3f} GUPS, BP {cbp:.3f} GUPS")

    others = None if args.no_other else other_workloads(ctx, stream, log)

    serve = None
    if not args.no_serve:
        serve = serving(args.serve_requests, args.serve_rate)

    line = {
        "metric": "cone-beam FP/BP GUPS 384^3x512",
        "value": round(value, 2),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"cone FDK: FP -> Hamming ramp filter -> BP, {W}^3 volume "
                               f"(cylinder r={140.0 * W / 384.0:g}, 0.02), {n} views "
                               "linspace(0,2pi), SAD 1000, detector WxH", "volume": [W, W, H],
                   "views": n, "geometry": "cone", "parallelism": "single GPU"},
        "fp_ms": round(fp_ms, 3), "filter_ms": round(fl_ms, 3), "bp_ms": round(bp_ms, 3),
        "fp_gups": round(fp_gups, 1), "bp_gups": round(bp_gups, 1),
        "roofline": roofline,
        "gather_roofline": gather,
        "fp16_storage": f16,
        "cpu_baseline": cpu,
        "other_workloads": others,
        "serving": serve,
    }
    print(json.dumps(line), flush=True)


def other_workloads(ctx, stream, log):
    """Device times of the other workloads: parallel-beam FP/BP of the 128^3
    phantom (180 views), helical general-3D FP/BP (320^3, 360 views of a
    256^2 jittered detector), and few-view OSC-TV (256^3, 60 views, 10
    subsets, 20 iterations, 10 TV steps) with FP32, FP16 and BF16 storage."""
    import torch

    import mexct
    from mexct import phantoms

    dev = torch.device("cuda:0")

    def dev_ms(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return round(e0.elapsed_time(e1) / reps, 3)

    res = {}
    log("[bench] other workloads: parallel 128^3 x 180")
    v = torch.from_numpy(phantoms.par_phantom_128(seed=1).T.copy()).to(dev)
    a = phantoms.angles_uniform(180, np.pi)
    p = torch.empty((180, 128, 128), dtype=torch.float32, device=dev)
    r = torch.empty((128, 128, 128), dtype=torch.float32, device=dev)
    res["parallel_128x180"] = {"fp_ms": dev_ms(lambda: ctx.fp(v, 1000.0, a, 180, 0, p)),
                               "bp_ms": dev_ms(lambda: ctx.bp(p, 1000.0, a, 0, r))}
    log("[bench] other workloads: 8 parallel 128^3 x 180 in one batched launch vs 8 launches")
    G = 8
    vb = v.unsqueeze(0).repeat(G, 1, 1, 1).contiguous()
    pb = torch.empty((G, 180, 128, 128), dtype=torch.float32, device=dev)
    rb = torch.empty((G, 128, 128, 128), dtype=torch.float32, device=dev)

    def fp_each():
        for g in range(G):
            ctx.fp(vb[g], 1000.0, a, 180, 0, pb[g])

    def bp_each():
        for g in range(G):
            ctx.bp(pb[g], 1000.0, a, 0, rb[g])
    res["batched_parallel_128x180x8"] = {
        "fp_batch_ms": dev_ms(lambda: ctx.fp_batch(vb, 1000.0, a, 180, 0, pb)),
        "fp_8_launches_ms": dev_ms(fp_each),
        "bp_batch_ms": dev_ms(lambda: ctx.bp_batch(pb, 1000.0, a, 0, rb)),
        "bp_8_launches_ms": dev_ms(bp_each)}
    del vb, pb, rb
    log("[bench] other workloads: helical general-3D 320^3 x 360 x 256^2")
    side, nv, wd = 320, 360, 256
    vol = phantoms.ellipsoid_phantom(side, side, 4, 12, 140.0, 0.02, (8.0, 40.0), (0.01, 0.03),
                                     100.0, 120.0)
    pts = phantoms.helix_pts(wd, wd, nv, side)
    vt = torch.from_numpy(vol.T.copy()).to(dev)
    pt = torch.from_numpy(np.ascontiguousarray(pts.transpose(3, 2, 1, 0))).to(dev)
    z = np.zeros(nv, np.float32)
    g = torch.empty((nv, wd, wd), dtype=torch.float32, device=dev)
    rv = torch.empty((side, side, side), dtype=torch.float32, device=dev)
    res["helix_g3d_320x360"] = {
        "fp_ms": dev_ms(lambda: ctx.g3d_fp(vt, z, pt, nv, g)),
        "fp16_fp_ms": dev_ms(lambda: ctx.g3d_fp(vt, z, pt, nv, g, storage=mexct.F16)),
        "bp_ms": dev_ms(lambda: ctx.g3d_bp(g, z, pt, side, side, rv), reps=3),
        "bp_deterministic_ms": dev_ms(lambda: ctx.g3d_bp(g, z, pt, side, side, rv, mode=1), reps=1),
        "bp_atomics_ms": dev_ms(lambda: ctx.g3d_bp(g, z, pt, side, side, rv, mode=2), reps=1),
        "bp_note": "bp_ms: voxel bricks in LDS (default); deterministic: the same with 64-bit "
                   "fixed-point sums; atomics: the ray-driven float-atomic scatter"}
    del vt, pt, g, rv
    log("[bench] other workloads: OSC-TV 256^3 x 60")
    W, n = 256, 60
    vol = phantoms.osc_phantom(W, seed=3)
    a = phantoms.angles_uniform(n, 2 * np.pi)
    fp = mexct.CUDAmex_FP(vol, 1000.0, a, n, 2)
    pre, post = phantoms.poisson_transmission(fp, 1e5, seed=33)
    picks = mexct.first_picks(303, 10, 10, 1.0, 20)
    pre_t = torch.from_numpy(pre.T.copy()).to(dev)
    post_t = torch.from_numpy(post.T.copy()).to(dev)
    mu = torch.empty((W, W, W), dtype=torch.float32, device=dev)
    osc = {}
    for name, st in (("f32", mexct.F32), ("f16", mexct.F16), ("bf16", mexct.BF16)):
        osc[name + "_ms"] = dev_ms(lambda st=st: ctx.osc_tv_recon(pre_t, post_t, a, 1000.0, 2, 20,
                                                                    0.1, 10, 10, 1.0, picks, mu,
                                                                    tv_steps=10, storage=st),
                                   reps=1)
    res["osc_tv_256x60"] = osc
    log(f"[bench] other workloads: {res}")
    return res


def same_geometry_burst(Server, phantoms, n=16, slots=2, max_group=8):
    """Bursts of n same-geometry requests (parallel 128^3, 180 views, f32)
    queued at once: servers that batch up to max_group of them per launch
    against servers that run them one by one (max_group 1), with host arrays
    (copied in and out) and with device-resident arrays (torch tensors, no
    host copies).  Each server first runs a burst untimed, which sizes its
    buffers and tunes the shapes; three more bursts are timed and the best
    one is reported beside all three (host bursts move by +-30% from burst
    to burst)."""
    import torch

    vol = phantoms.par_phantom_128(seed=1)
    ang = phantoms.angles_uniform(180, np.pi)
    sino = np.asfortranarray(np.full((128, 128, 180), 0.5, np.float32))
    dvol = torch.from_numpy(np.ascontiguousarray(vol.T)).cuda()
    dsino = torch.from_numpy(np.ascontiguousarray(sino.T)).cuda()
    out = {"workload": f"bursts of {n} requests, parallel 128^3, 180 views, f32, queued at once; "
                       "three bursts on a warm server are timed, the best reported; host "
                       "arrays, or resident device arrays",
           "slots": slots}
    for kind in ("fp", "bp"):
        for resident in (False, True):
            v, p = (dvol, dsino) if resident else (vol, sino)
            for mg in (1, max_group):
                with Server(hbm_budget=16 << 30, slots=slots, max_group=mg) as s:
                    sub = ((lambda: s.submit_fp(v, 1000.0, ang, 180, 0)) if kind == "fp"
                           else (lambda: s.submit_bp(p, 1000.0, ang, 0)))
                    for t in [sub() for _ in range(n)]:
                        s.result(t)
                    b0 = s.stats()["batched"]
                    rates = []
                    for _ in range(3):
                        t0 = time.perf_counter()
                        for t in [sub() for _ in range(n)]:
                            s.result(t)
                        rates.append(n / (time.perf_counter() - t0))
                    st = s.stats()
                key = f"{kind}_{'resident_' if resident else ''}max_group_{mg}"
                out[key] = {"throughput_req_s": round(max(rates), 1),
                            "bursts_req_s": [round(r, 1) for r in rates],
                            "batched": st["batched"] - b0}
    return out


SERVE_SIDES = (128, 192, 256, 384, 512)
SERVE_VIEWS = (90, 180, 360, 720)


def serving_stream(n_req, rate, seed=5):
    """The mixed request stream: 40% FP, 30% BP, 20% FDK, 10% OSC-TV (5
    iterations), cubic sides from SERVE_SIDES, views from SERVE_VIEWS,
    par/fan/cone, SAD ~ U[800,1200], FP32 or FP16 storage, Poisson arrivals
    at `rate` per second."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kinds = rng.choice(["fp", "bp", "fdk", "osc"], size=n_req, p=[0.4, 0.3, 0.2, 0.1])
    reqs = []
    t = 0.0
    for k in kinds:
        t += float(rng.exponential(1.0 / rate))
        reqs.append({"kind": str(k), "side": int(rng.choice(SERVE_SIDES)),
                     "views": int(rng.choice(SERVE_VIEWS)), "geom": int(rng.integers(0, 3)),
                     "sad": float(np.float32(rng.uniform(800.0, 1200.0))),
                     "storage": int(rng.integers(0, 2)), "arrival": t})
    return reqs


def serving(n_req, rate):
    """Open-loop serving run of serving_stream through mexct.serve.Server,
    a saturated run (all requests queued at once) and the one-at-a-time
    plain-FP32 loop over the same requests through the direct entry points.
    Latency = arrival -> output in host memory (host arrays in and out)."""
    import threading

    import mexct
    from mexct import phantoms
    from mexct.serve import Server

    reqs = serving_stream(n_req, rate)
    log(f"[serve] {n_req} requests, rate {rate}/s; building inputs")
    vols, sinos, angs = {}, {}, {}
    for r in reqs:
        W, v = r["side"], r["views"]
        if W not in vols:
            vols[W] = phantoms.cylinder(W, W, 0.36 * W)
        if (W, v) not in sinos:
            sinos[(W, v)] = np.asfortranarray(np.full((W, W, v), 0.5, np.float32))
        if v not in angs:
            angs[v] = phantoms.angles_uniform(v, 2 * np.pi)
    pre_c = {}

    def osc_args(r):
        W, v = r["side"], r["views"]
        if (W, v) not in pre_c:
            pre_c[(W, v)] = (np.asfortranarray(np.full((W, W, v), 1e5, np.float32)),
                             np.asfortranarray(np.full((W, W, v), 6e4, np.float32)))
        pre, post = pre_c[(W, v)]
        g = {"type": ["par3d", "fan3d", "cone3d"][r["geom"]], "SAD": r["sad"]}
        return pre, post, angs[v], g, 5, 0.1, 10, 10, 1.0

    def submit(s, r):
        W, v, g, sad, st = r["side"], r["views"], r["geom"], r["sad"], r["storage"]
        if r["kind"] == "fp":
            return s.submit_fp(vols[W], sad, angs[v], v, g, storage=st)
        if r["kind"] == "bp":
            return s.submit_bp(sinos[(W, v)], sad, angs[v], g, storage=st)
        if r["kind"] == "fdk":
            return s.submit_fdk(sinos[(W, v)], {"type": ["par3d", "fan3d", "cone3d"][g],
                                                "SAD": sad}, angs[v], "hamming", 1.0,
                                storage=st)
        return s.submit_osc_tv(*osc_args(r), tv_steps=10, picks=np.ones(5, np.int32), storage=st)

    def run_alone(r):
        W, v, g, sad = r["side"], r["views"], r["geom"], r["sad"]
        if r["kind"] == "fp":
            return mexct.CUDAmex_FP(vols[W], sad, angs[v], v, g)
        if r["kind"] == "bp":
            return mexct.CUDAmex_BP(sinos[(W, v)], sad, angs[v], g)
        if r["kind"] == "fdk":
            geo = {"type": ["par3d", "fan3d", "cone3d"][g], "SAD": sad}
            return mexct.CUDAmex_BP(mexct.preFilterSinogram(sinos[(W, v)], geo, angs[v],
                                                            "hamming", 1.0), sad, angs[v], g)
        return mexct.OSC_TV_recon(*osc_args(r), tv_steps=10, picks=np.ones(5, np.int32))

    budget = 96 << 30

    def served(open_loop, max_group=4):
        lats, queues = [None] * n_req, [None] * n_req
        with Server(hbm_budget=budget, slots=4, max_group=max_group) as s:
            tickets = [None] * n_req
            ready = threading.Semaphore(0)

            def waiter():
                for i in range(n_req):
                    ready.acquire()
                    _, lat, q = s.wait(tickets[i][0])
                    lats[i] = tickets[i][1] + lat
                    queues[i] = q

            th = threading.Thread(target=waiter)
            th.start()
            t0 = time.perf_counter()
            for i, r in enumerate(reqs):
                if open_loop:
                    d = t0 + r["arrival"] - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    lag = time.perf_counter() - (t0 + r["arrival"])
                else:
                    lag = 0.0
                tickets[i] = (submit(s, r), lag)
                ready.release()
            th.join()
            wall = time.perf_counter() - t0
            st = s.stats()
        return np.array(lats), wall, st

    log("[serve] saturated run")
    lat_sat, wall_sat, st_sat = served(False)
    log(f"[serve] saturated: {n_req / wall_sat:.2f} req/s, stats {st_sat}")
    log("[serve] open-loop run")
    lat_open, wall_open, st_open = served(True)
    log(f"[serve] open loop: p50 {np.percentile(lat_open, 50) * 1e3:.1f} ms "
        f"p99 {np.percentile(lat_open, 99) * 1e3:.1f} ms")
    log("[serve] open-loop run without batching (max_group 1)")
    lat_nob, wall_nob, st_nob = served(True, max_group=1)
    log(f"[serve] open loop, no batching: p50 {np.percentile(lat_nob, 50) * 1e3:.1f} ms "
        f"p99 {np.percentile(lat_nob, 99) * 1e3:.1f} ms")
    log("[serve] same-geometry burst")
    burst = same_geometry_burst(Server, phantoms)
    log(f"[serve] burst: {burst}")
    log("[serve] one-at-a-time FP32 loop")
    svc = []
    for r in reqs:
        t = time.perf_counter()
        run_alone(r)
        svc.append(time.perf_counter() - t)
    svc = np.array(svc)
    # the same arrivals through a FIFO single server with these service times
    arr = np.array([r["arrival"] for r in reqs])
    done = np.zeros(n_req)
    prev = 0.0
    for i in range(n_req):
        prev = max(arr[i], prev) + svc[i]
        done[i] = prev
    lat_seq = done - arr
    pct = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 2)  # noqa: E731
    return {
        "workload": f"{n_req} requests (seed 5): 40% FP, 30% BP, 20% FDK, 10% OSC-TV "
                    "(5 iterations, 10 subsets, 10 TV steps); cubic sides "
                    f"{list(SERVE_SIDES)}, views {list(SERVE_VIEWS)}, par/fan/cone, "
                    "SAD U[800,1200], f32 or f16 storage; host arrays in and out",
        "arrival_rate_per_s": rate,
        "hbm_budget_bytes": budget, "slots": 4, "max_group": 4,
        "open_loop": {"p50_ms": pct(lat_open, 50), "p99_ms": pct(lat_open, 99),
                      "mean_ms": round(float(lat_open.mean()) * 1e3, 2),
                      "throughput_req_s": round(n_req / wall_open, 3),
                      "peak_reserved_gb": round(st_open["peak_reserved_bytes"] / 1e9, 2),
                      "batched": st_open["batched"]},
        "open_loop_no_batching": {"max_group": 1, "p50_ms": pct(lat_nob, 50),
                                  "p99_ms": pct(lat_nob, 99),
                                  "mean_ms": round(float(lat_nob.mean()) * 1e3, 2),
                                  "throughput_req_s": round(n_req / wall_nob, 3),
                                  "note": "every request of this stream has its own SAD, so no "
                                          "two share a batch key; the same-geometry bursts "
                                          "measure batching"},
        "saturated": {"throughput_req_s": round(n_req / wall_sat, 3),
                      "p50_ms": pct(lat_sat, 50), "p99_ms": pct(lat_sat, 99),
                      "groups": st_sat["groups"], "grouped": st_sat["grouped"],
                      "batched": st_sat["batched"]},
        "same_geometry_burst": burst,
        "one_at_a_time_fp32": {"throughput_req_s": round(n_req / float(svc.sum()), 3),
                               "p50_ms": pct(lat_seq, 50), "p99_ms": pct(lat_seq, 99),
                               "note": "measured service times replayed through a FIFO "
                                       "single server with the same arrivals"},
    }


if __name__ == "__main__":
    main()
