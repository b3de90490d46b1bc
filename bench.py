#!/usr/bin/env python3
"""Benchmark of the MI355X film-grain hot path.

Default workload (the headline metric): pixel-wise render of a 1920x1080 RGB
frame at zoom 2 (3840x2160 output), lognormal radii mean 0.1 / sd 0.05, filter
sigma 0.8, N = 256, seed 7, with the lambda planes already resident in HBM.
One step renders the three colour planes.  The value is
Mpix*samples/s = 3840 * 2160 * 3 * 256 * steps / seconds / 1e6.

Also reported on the same JSON line (default workload, after its timed region):
  roofline      FP32 FLOP/s of the pixel-wise splat kernel against MI355X's
                FP32 peak; FLOPs per (pixel, sample) = 16 + 6*T, T = disk tests
                in the reference's loop order with its early exit, counted by
                the kernel's instrumented twin on the same frame; duration per
                launch from HIP events the library records on the launch
                stream, over the frame's calls as the timed steps make them
                (one persistent splat launch covers the frame's planes).
  cpu_baseline  the CPU oracle (a C restatement of the reference's CPU path)
                on a bounded random sample of 16x16 output tiles of the same
                frame, on the host's cores; those tiles also spot-check parity.

  grain_c3      the grain-wise 4096x4096 C3 frame (N = 128): Mpix*samples/s,
                its HBM roofline (8 + N/4 algorithmic bytes per output pixel)
                and a VALU roofline of the reference's own work (8 FLOPs per
                (grain, sample) pair plus 5 per disk test, counted exactly)
  batch_c4      the uneven 64-image batch in one graph-captured launch
                against one call per plane
  frame         (default workload) the frame from host memory to host memory
                (PCIe included): host split / lambda / merge against the
                frame on the device, with identical output.
  adaptive      (default workload) the learned per-pixel sample budget on a
                1920x1080 RGB frame: adaptive ms/frame at the RMSE fixed
                N = 1024 reaches against the N = 4096 render, fixed-N ms/frame
                at the same RMSE, allocators at equal sample totals, and the
                error model's BF16 MFMA throughput (after the timed region).

    python bench.py [--steps K] [--warmup W]
                    [--workload pixel-c2|grain-c3|pixel-c1|batch-c4|adaptive-c5]
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

from film_grain_amd.workloads import WORKLOADS, prepare  # noqa: E402

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_TABLE = os.path.join(ROOT, "film_grain_amd", "tuning", "traffic.json")


def traffic_key(workload, kernel, cfg):
    launch = ",".join(f"{k}={cfg[k]}" for k in sorted(cfg) if cfg[k])
    return f"{workload}|{kernel}|{launch}"


def measured_traffic(workload, kernel, cfg):
    """HBM bytes per launch of the dominant kernel, from the FETCH_SIZE and
    WRITE_SIZE counter runs recorded for this workload and launch shape
    (film_grain_amd.profiling); None when none were recorded."""
    try:
        with open(TRAFFIC_TABLE) as f:
            rec = json.load(f).get(traffic_key(workload, kernel, cfg))
    except (OSError, ValueError):
        return None
    return None if rec is None else rec["traffic_bytes"]


def _cpu_model() -> str:
    """The host CPU's model name (lscpu, else /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads() -> int:
    """Every core this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(wl, planes, lam0, derived, offs, gpu_plane0, budget_s=15.0):
    """Times the oracle on random 16x16 output tiles of plane 0 and checks the
    GPU frame on the same tiles; also a one-thread rate on a smaller slice of
    the same kind of sample, and the host's CPU model."""
    from oracle import oracle as O
    threads = _host_threads()
    od = O.derive(wl["radius"], wl["sd"], wl["dist"], wl["zoom"], 0.8, wl["iters"], seed=wl["seed"],
                  in_w=wl["w"], in_h=wl["h"])
    p = O.plane(od, wl["iters"], wl["seed"], wl["zoom"], wl["dist"], wl["radius"])
    rng = np.random.default_rng(2024)
    tile = 16
    n_px = 0
    mism = 0
    t_used = 0.0
    n_tiles = 0
    if wl["algo"] == "grain":
        # the grain-wise oracle renders whole frames: time random 256x256
        # crops of the input (the same lambda values, so the same work per
        # pixel; crops do not reproduce the full frame's random streams, so
        # they are timed only)
        crop = min(wl["w"], wl["h"], 256)
        od_c = O.derive(wl["radius"], wl["sd"], wl["dist"], wl["zoom"], 0.8, wl["iters"],
                        seed=wl["seed"], in_w=crop, in_h=crop)
        pc = O.plane(od_c, wl["iters"], wl["seed"], wl["zoom"], wl["dist"], wl["radius"])
        n_crops = 0
        first = None
        while t_used < budget_s and n_crops < 64:
            x0 = int(rng.integers(0, wl["w"] - crop + 1))
            y0 = int(rng.integers(0, wl["h"] - crop + 1))
            lam_c = np.ascontiguousarray(lam0[y0:y0 + crop, x0:x0 + crop])
            if first is None:
                first = lam_c
            t0 = time.perf_counter()
            O.render_grainwise(pc, lam_c, offs, threads=threads)
            t_used += time.perf_counter() - t0
            n_px += pc.out_w * pc.out_h
            n_crops += 1
        sample = (f"grain-wise oracle on {n_crops} random {crop}x{crop} input crops of the frame "
                  f"({n_px * wl['iters']:.3g} pixel-samples, {t_used:.1f} s)")
        t0 = time.perf_counter()
        O.render_grainwise(pc, first, offs, threads=1)
        one = pc.out_w * pc.out_h * wl["iters"] / (time.perf_counter() - t0) / 1e6
        return dict(value=round(n_px * wl["iters"] / t_used / 1e6, 3), unit="Mpix·samples/s", cores=threads,
                    kind="port", sample=sample, one_thread_value=round(one, 3),
                    one_thread_sample="the first of those crops", cpu_model=_cpu_model(),
                    nproc=os.cpu_count()), None
    while t_used < budget_s:
        x0 = int(rng.integers(0, od.out_w - tile + 1))
        y0 = int(rng.integers(0, od.out_h - tile + 1))
        t0 = time.perf_counter()
        want = O.render_pixelwise(p, lam0, offs, region=(x0, y0, x0 + tile, y0 + tile), threads=threads)
        t_used += time.perf_counter() - t0
        n_px += tile * tile
        n_tiles += 1
        got = gpu_plane0[y0:y0 + tile, x0:x0 + tile]
        mism += int(np.sum(got != want))
        if n_tiles >= 400:
            break
    value = n_px * wl["iters"] / t_used / 1e6
    sample = (f"{n_tiles} random {tile}x{tile} output tiles of colour plane 0 "
              f"({n_px * wl['iters']:.3g} pixel-samples, {t_used:.1f} s)")
    # one thread, on the first 24 tiles of the same random sequence
    rng1 = np.random.default_rng(2024)
    n1, t1 = 0, 0.0
    for _ in range(min(24, n_tiles)):
        x0 = int(rng1.integers(0, od.out_w - tile + 1))
        y0 = int(rng1.integers(0, od.out_h - tile + 1))
        t0 = time.perf_counter()
        O.render_pixelwise(p, lam0, offs, region=(x0, y0, x0 + tile, y0 + tile), threads=1)
        t1 += time.perf_counter() - t0
        n1 += tile * tile
    one = n1 * wl["iters"] / t1 / 1e6 if t1 > 0 else None
    return dict(value=round(value, 3), unit="Mpix·samples/s", cores=threads, kind="port",
                sample=sample, one_thread_value=None if one is None else round(one, 3),
                one_thread_sample=f"the first {min(24, n_tiles)} of those tiles", cpu_model=_cpu_model(),
                nproc=os.cpu_count()), \
        dict(tiles=n_tiles, pixels=n_px, mismatched_pixels=mism)


def batch_summary(steps: int, warmup: int, n_images: int = 64, tuning: bool = True):
    """The uneven batch: 64 images (sides 128-2048, mixed radii, N and
    channels), every plane resident in HBM; one graph-captured batch run per
    step, against the same planes rendered one call at a time."""
    import torch
    from film_grain_amd import capi, workloads
    images = workloads.batch_images(master_seed=2024, n_images=n_images)
    planes = workloads.prepare_batch(images)
    dev = torch.device("cuda:0")
    ins = [(d, torch.from_numpy(l).to(dev), torch.from_numpy(o).to(dev)) for d, l, o in planes]
    outs = [torch.empty((d.out_h, d.out_w), dtype=torch.float32, device=dev) for d, _, _ in ins]
    px_samples = sum(d.out_w * d.out_h * d.n_samples for d, _, _ in ins)
    log(f"[bench] batch: {len(images)} images, {len(ins)} planes, {px_samples:.4g} pixel-samples")
    ctx = capi.Context(0)
    if not tuning:
        ctx.set_launch_table([])
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    batch = capi.Batch(ctx, [(d, l, o, out) for (d, l, o), out in zip(ins, outs)], capture_graph=True)
    plan_s = time.perf_counter() - t0
    log(f"[bench] batch planned in {plan_s:.2f} s: {batch.info()}")
    for _ in range(warmup):
        batch.run(stream)
    batch.status(stream)
    ref = [o.clone() for o in outs]

    def timed(fn, n):
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / n

    batch_s = timed(lambda: batch.run(stream), steps)
    batch.status(stream)

    def per_plane():
        for (d, l, o), out in zip(ins, outs):
            ctx.render_device("pixel", d, l, o, out, stream)

    per_plane()
    ctx.stream_status(stream)
    single_s = timed(per_plane, max(1, steps // 2))
    ctx.stream_status(stream)
    same = all(torch.equal(a, b) for a, b in zip(outs, ref))
    log(f"[bench] batch {batch_s * 1e3:.2f} ms, one call per plane {single_s * 1e3:.2f} ms, identical planes: {same}")
    info = batch.info()
    batch.close()
    ctx.close()
    return {
        "metric": "Mpix·samples/s (uneven batch, one graph-captured persistent launch)",
        "value": round(px_samples / batch_s / 1e6, 2), "unit": "Mpix·samples/s", "steps": steps,
        "ms_per_step": round(batch_s * 1e3, 3),
        "workload": f"{len(images)} images, sides 128-2048, radius U(0.05,0.3), sigma_r/r U(0,0.5), "
                    f"N in {{32..512}}, luma/RGB, seed 2024",
        "planes": len(ins), "pixel_samples": px_samples, "plan_s": round(plan_s, 3), "info": info,
        "per_plane_calls_ms": round(single_s * 1e3, 3),
        "speedup_vs_per_plane_calls": round(single_s / batch_s, 3),
        "identical_to_per_plane_calls": bool(same),
    }


# Algorithmic FLOPs per row of the error model (12 inputs, 64, 64, 1): the
# forward pass, and forward + backward (dW3, dZ2, dW2, dH1, dW1).
MLP_FWD_FLOPS = 2 * (12 * 64 + 64 * 64 + 64)
MLP_GRAD_FLOPS = MLP_FWD_FLOPS + 2 * 64 + 2 * 64 + 2 * (2 * 64 * 64) + 2 * 12 * 64
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def frame_summary(reps: int = 3):
    """The whole C2 frame from host memory to host memory (1920x1080 RGB f32
    in, 3840x2160 8-bit RGB out, PCIe included): the host path (split, lambda
    and merge on the host, one host-memory integrator call per plane) against
    the frame on the device (one upload, split, lambda, integrators, merge,
    one download).  Outputs must be identical."""
    from film_grain_amd import host
    from film_grain_amd.workloads import make_planes
    wl = WORKLOADS["pixel-c2"]
    rgb = np.ascontiguousarray(np.transpose(make_planes("pixel-c2"), (1, 2, 0)))
    args = host.CliArgs(radius=wl["radius"], radius_stddev=wl["sd"], radius_dist=wl["dist"], zoom=wl["zoom"],
                        iters=wl["iters"], seed=wl["seed"], algo="pixel", sigma=0.8, color_mode="rgb")
    out = {}
    res = {}
    for name, on_dev in (("host_path_ms", False), ("frame_on_device_ms", True)):
        host.render_rgb(args, rgb, frame_on_device=on_dev, with_planes=False)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            res[name], _, _ = host.render_rgb(args, rgb, frame_on_device=on_dev, with_planes=False)
            ts.append((time.perf_counter() - t0) * 1e3)
        out[name] = round(float(np.median(ts)), 3)
    out["identical"] = bool(np.array_equal(res["host_path_ms"], res["frame_on_device_ms"]))
    out["workload"] = "pixel-c2 frame: 1920x1080 RGB f32 in host memory -> 3840x2160 RGB8 in host memory"
    log(f"[bench] frame from host memory: host path {out['host_path_ms']} ms, on the device "
        f"{out['frame_on_device_ms']} ms, identical {out['identical']}")
    return out


def adaptive_summary(steps: int = 3, warmup: int = 1):
    """The learned sample budget on one 1920x1080 RGB frame (lognormal
    0.1/0.05, sigma 0.8, zoom 1, budgets 16..1024).  The headline allocator
    gives each tile of the adaptive renderer one sample level, chosen from
    the error model's per-pixel spreads and the tile's measured per-sample
    cost (n ~ s / sqrt(c)); it is timed whole (features, error model,
    levels and renders) at the RMSE that fixed N = 256, 512 and 1024 reach
    against the N = 4096 render, beside the fixed-N ms/frame at the adaptive
    frame's own RMSE.  Also: the per-pixel allocator, the allocators at
    equal sample totals, and the error model's MFMA throughput."""
    import torch
    from film_grain_amd import budget, capi
    ctx = capi.Context(0)
    model = budget.ErrorModel.load()
    specs = budget.image_planes(50_000)
    cal = budget.calibrate_cost(ctx, specs[0])
    log(f"[bench] cost calibration: kappa {cal['kappa']:.4g} (ms per sample {cal['ms_per_sample']}, "
        f"lambda {cal['lambdas']})")
    planes = [budget.upload(sp) for sp in specs]
    refs = [budget.render_reference(ctx, dp) for dp in planes]
    curve = budget.fixed_curve(ctx, planes, refs, ns=(128, 256, 512, 1024, 2048), reps=steps)
    bufs = budget.frame_buffers(planes)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ts = []
        res = None
        for _ in range(steps):
            t0 = time.perf_counter()
            res = fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts)), res

    def mean_budget():
        return float(np.mean([float((b.to(torch.int32) & 0xFFFF).float().mean()) for b in bufs.budget]))

    targets = []
    for tn in (256, 512, 1024):
        target = next(e for n, ms_, e in curve if n == tn)
        fixed_tn = next(ms_ for n, ms_, e in curve if n == tn)
        row = dict(target_fixed_n=tn, target_rmse_255=round(target * 255, 4), fixed_ms=round(fixed_tn, 3))
        runs = {
            "per_pixel": lambda: budget.render_adaptive_frame(ctx, model, planes, bufs, target ** 2, "learned"),
            "tile_levels": lambda: budget.render_tiled_frame(ctx, model, planes, bufs, target ** 2, cal["kappa"],
                                                             "learned", refs=refs),
            "tile_levels_oracle": lambda: budget.render_tiled_frame(ctx, model, planes, bufs, target ** 2,
                                                                    cal["kappa"], "oracle", refs=refs),
        }
        for kind, fn in runs.items():
            ms_, (outs, _info) = timed(fn)
            rmse = budget._frame_rmse(outs, refs)
            fx = budget.ms_at_rmse(curve, rmse)
            row[kind] = dict(ms=round(ms_, 3), rmse_255=round(rmse * 255, 4), mean_budget=round(mean_budget(), 1),
                             fixed_ms_at_rmse=None if fx is None else round(fx, 3),
                             fixed_over_adaptive=None if fx is None else round(fx / ms_, 4))
            log(f"[bench] adaptive ({kind}) at fixed N={tn}'s rmse {target * 255:.3f}/255: "
                f"{ms_:.2f} ms/frame rmse {rmse * 255:.3f}/255 mean n {mean_budget():.1f}; fixed at that rmse "
                f"{fx if fx is None else round(fx, 2)} ms (fixed N={tn}: {fixed_tn:.2f} ms)")
        targets.append(row)
    # the headline: the per-pixel allocator (the device scale search, no host
    # round trip) at the RMSE of fixed N = 1024
    head = targets[-1]["per_pixel"]
    ms, rmse, mean_n, fixed_ms = head["ms"], head["rmse_255"] / 255, head["mean_budget"], head["fixed_ms_at_rmse"]
    fixed_1024 = next(ms_ for n, ms_, e in curve if n == budget.C5["n_fixed"])
    ratio = budget.s_ratio(refs)
    eq = budget.compare_at_equal_budget(ctx, model, budget.image_planes(50_000), mean_budgets=(256,))
    # the error model's kernels on the frame's features
    X = torch.cat(bufs.X)
    rows = X.shape[0]
    pred = torch.empty(rows, dtype=torch.float32, device=X.device)
    grad = torch.zeros(capi.MLP_PARAMS, dtype=torch.float32, device=X.device)
    loss = torch.zeros(1, dtype=torch.float32, device=X.device)
    e = torch.zeros(rows, dtype=torch.float32, device=X.device)
    ctx.mlp_forward(model.P, X, out=pred)
    ctx.mlp_grad(model.P, X, e, grad, loss, loss_kind="nll")
    st = torch.cuda.current_stream()
    a, b, g0, g1 = (torch.cuda.Event(enable_timing=True) for _ in range(4))
    a.record(st)
    for _ in range(5):
        ctx.mlp_forward(model.P, X, out=pred)
    b.record(st)
    g0.record(st)
    for _ in range(5):
        ctx.mlp_grad(model.P, X, e, grad, loss, loss_kind="nll")
    g1.record(st)
    torch.cuda.synchronize()
    fwd_ms, grad_ms = a.elapsed_time(b) / 5, g0.elapsed_time(g1) / 5
    fwd_tf = rows * MLP_FWD_FLOPS / (fwd_ms * 1e-3) / 1e12
    grad_tf = rows * MLP_GRAD_FLOPS / (grad_ms * 1e-3) / 1e12
    log(f"[bench] error model on {rows} rows: forward {fwd_ms:.3f} ms ({fwd_tf:.1f} TFLOP/s), "
        f"forward+backward {grad_ms:.3f} ms ({grad_tf:.1f} TFLOP/s)")
    ctx.close()
    return {
        "metric": "ms/frame adaptive vs fixed-N at equal RMSE",
        "workload": "1920x1080 RGB, lognormal 0.1/0.05, sigma 0.8, zoom 1, budgets 16..1024, seed 50000; "
                    "RMSE against the N=4096 render; targets = the RMSE of fixed N=256, 512, 1024; "
                    "headline at the N=1024 target",
        "allocator": "per-pixel budgets ceil(c s) from the error model's spreads (headline); tile levels with the "
                     "measured per-sample cost (n ~ s/sqrt(c)) reported per target",
        "adaptive_ms_per_frame": round(ms, 3), "rmse_255": round(rmse * 255, 4), "mean_budget": round(mean_n, 1),
        "fixed_ms_at_equal_rmse": fixed_ms,
        "fixed_over_adaptive": None if fixed_ms is None else round(fixed_ms / ms, 4),
        "fixed_n1024_ms_per_frame": round(fixed_1024, 3),
        "targets": targets,
        "cost_model": {"kappa": round(cal["kappa"], 6), "ms_per_sample": [round(v, 6) for v in cal["ms_per_sample"]],
                       "lambdas": [round(v, 3) for v in cal["lambdas"]]},
        "fixed_curve": [dict(n=n, ms=round(t, 3), rmse_255=round(e_ * 255, 4)) for n, t, e_ in curve],
        "allocation_bound_E_s2_over_E_s_sq": round(ratio, 4),
        "equal_samples_256_rmse_255": {r["kind"]: round(r["rmse"] * 255, 4) for r in eq},
        "mlp": {"rows": rows, "forward_ms": round(fwd_ms, 4), "forward_tflops": round(fwd_tf, 2),
                "grad_ms": round(grad_ms, 4), "grad_tflops": round(grad_tf, 2), "peak_bf16_tflops": PEAK_BF16_TFLOPS,
                "flops_per_row": {"forward": MLP_FWD_FLOPS, "forward_backward": MLP_GRAD_FLOPS}},
    }


def measure(workload: str, steps: int, warmup: int, cpu: bool, cpu_seconds: float, stats: bool = True,
            tuning: bool = True, pix_variant=None, copy_calibration_mib: int = 0):
    """One workload of WORKLOADS on one GPU: the timed steps, the kernel
    timing and roofline, and the CPU baseline (with its parity spot check)."""
    import torch
    from film_grain_amd import autotune, capi

    wl = WORKLOADS[workload]
    log(f"[bench] workload {workload}: {wl['desc']}")
    prep = prepare(workload)
    planes, lams, offsets, d, desc, algo = prep.planes, prep.lams, prep.offsets, prep.derived, prep.desc, prep.algo

    dev = torch.device("cuda:0")
    ctx = capi.Context(0)
    lam_d = [torch.from_numpy(l).to(dev) for l in lams]
    offs_d = torch.from_numpy(np.ascontiguousarray(offsets)).to(dev)
    out_d = torch.empty((len(lams), d.out_h, d.out_w), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    if copy_calibration_mib > 0:
        n = copy_calibration_mib << 18  # float32 elements
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {copy_calibration_mib} MiB done")
        del src, dst

    # instrumented pass: algorithmic work of the frame
    counters = []
    if stats:
        for i, l in enumerate(lam_d):
            counters.append(ctx.stats_device(algo, desc, l, offs_d, out_d[i], stream))
        torch.cuda.synchronize()
        log(f"[bench] counters per plane {counters}")

    # the library's built-in tuned table serves the render unless disabled
    if not tuning:
        ctx.set_launch_table([])
    key = autotune.bucket_key(algo, desc)
    tuned = tuning and any(capi.bucket_name(k) == key for k, _ in ctx.launch_table())
    if pix_variant is not None:
        ctx.set_launch_config(pix_variant=pix_variant)
        tuned = False
    cfg = ctx.launch_config(desc, algo)
    log(f"[bench] launch config {cfg} (bucket {key}, tuned table entry: {tuned})")

    # a frame's pixel-wise planes go through one call, which hashes each
    # cell's stream once for all of them (the planes share the seed)
    joint = algo == "pixel" and 1 < len(lam_d) <= 3 and pix_variant is None
    outs = [out_d[i] for i in range(len(lam_d))]

    def step_frame():
        if joint:
            ctx.render_planes_device(desc, lam_d, offs_d, outs, stream)
        else:
            for i, l in enumerate(lam_d):
                ctx.render_device(algo, desc, l, offs_d, out_d[i], stream)

    for _ in range(warmup):
        step_frame()
    ctx.stream_status(stream)
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        step_frame()
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.stream_status(stream)
    secs = t1 - t0
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    log(f"[bench] {steps} steps in {secs:.4f} s; step ms mean {float(np.mean(launch_ms)):.3f} "
        f"min {min(launch_ms):.3f} max {max(launch_ms):.3f} ({'one call for the frame' if joint else 'one call per plane'})")

    # kernel timing pass: HIP events recorded by the library on the launch stream
    # around the generation kernels and around the coverage-test kernel
    ctx.set_timing(True)
    kt = []
    for _ in range(max(2, steps)):
        for i, l in enumerate(lam_d):
            ctx.render_device(algo, desc, l, offs_d, out_d[i], stream)
            kt.append(ctx.timings())
    ctx.set_timing(False)
    ctx.stream_status(stream)
    kt = np.array(kt)
    render_ms, gen_ms, test_ms = (float(v) for v in kt.mean(axis=0))
    log(f"[bench] per plane: render {render_ms:.3f} ms = generation {gen_ms:.3f} + test kernel {test_ms:.3f}")
    # the frame's call, as the timed steps run it: its planes' splats are one
    # persistent launch (splat_batch_kernel), timed by the library's events
    frame_kt = None
    if joint:
        ctx.set_timing(True)
        ft = []
        for _ in range(max(2, steps)):
            ctx.render_planes_device(desc, lam_d, offs_d, outs, stream)
            ft.append(ctx.timings())
        ctx.set_timing(False)
        ctx.stream_status(stream)
        frame_kt = [float(v) for v in np.array(ft).mean(axis=0)]
        log(f"[bench] frame: render {frame_kt[0]:.3f} ms = generation {frame_kt[1]:.3f} + "
            f"splat kernel (all planes) {frame_kt[2]:.3f}")

    px_samples = d.out_w * d.out_h * len(lam_d) * d.n_samples
    value = px_samples * steps / secs / 1e6
    kernel_name = {1: "pixelwise_gather", 2: "pixelwise_splat", 3: "splat_hbm_kernel"}.get(cfg.get("pix_variant"), "")
    if frame_kt is not None:
        kernel_name = "splat_batch_kernel"
    if algo == "grain":
        kernel_name = "grainwise_sorted" if cfg.get("grain_variant") == 2 else "grainwise_tiles"
    traffic = measured_traffic(workload, kernel_name, cfg)
    log(f"[bench] traffic key {traffic_key(workload, kernel_name, cfg)} -> {traffic}")
    valu = None
    if not counters:
        roofline = None
    elif algo == "pixel":
        flops = [16.0 * c[0] + 6.0 * c[1] for c in counters]
        # one launch splats every plane of the frame, or one plane
        launch_flops = float(np.sum(flops)) if frame_kt is not None else float(np.mean(flops))
        k_ms, k_gen, k_render = (frame_kt[2], frame_kt[1], frame_kt[0]) if frame_kt is not None else \
            (test_ms, gen_ms, render_ms)
        achieved = launch_flops / (k_ms * 1e-3) / 1e12
        roofline = dict(bound="valu", achieved=round(achieved, 3), peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                        frac=round(achieved / PEAK_FP32_TFLOPS, 5), traffic=traffic,
                        kernel=kernel_name, kernel_ms=round(k_ms, 3),
                        generation_ms=round(k_gen, 3), render_ms=round(k_render, 3),
                        flops_per_launch=launch_flops, planes_per_launch=len(flops) if frame_kt is not None else 1,
                        tests_per_point=round(sum(c[1] for c in counters) / sum(c[0] for c in counters), 4))
    else:
        # algorithmic bytes per output pixel: lambda read 4 + output write 4 + N/4 (bitset clear + read)
        bytes_per_launch = d.out_w * d.out_h * (8 + d.n_samples / 4.0)
        achieved = bytes_per_launch / (test_ms * 1e-3) / 1e9
        roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=PEAK_HBM_GBS, unit="GB/s",
                        frac=round(achieved / PEAK_HBM_GBS, 6), traffic=traffic,
                        kernel=kernel_name, kernel_ms=round(test_ms, 3), generation_ms=round(gen_ms, 3),
                        render_ms=round(render_ms, 3), bytes_per_launch=bytes_per_launch)
        # the reference's own work: 8 FLOPs per (grain, sample) pair for its
        # bounds and 5 per disk test, counted by the instrumented twin
        pairs = float(np.mean([c[1] for c in counters]))
        tests = float(np.mean([c[2] for c in counters]))
        vflops = 8.0 * pairs + 5.0 * tests
        va = vflops / (test_ms * 1e-3) / 1e12
        valu = dict(bound="valu", achieved=round(va, 3), peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=round(va / PEAK_FP32_TFLOPS, 5), flops_per_launch=vflops, pairs_per_launch=pairs,
                    disk_tests_per_launch=tests, grains_per_launch=float(np.mean([c[0] for c in counters])))
        log(f"[bench] grain-wise work: {pairs:.4g} pairs, {tests:.4g} disk tests -> {va:.2f} TFLOP/s")

    cpu_b = None
    parity = None
    if cpu:
        plane0 = out_d[0].cpu().numpy()
        cpu_b, parity = cpu_baseline(wl, planes, lams[0], d, offsets, plane0, cpu_seconds)
        log(f"[bench] cpu baseline {cpu_b} parity {parity}")
    ctx.close()
    line = {
        "metric": "Mpix·samples/s at 4K RGB N=256" if workload == "pixel-c2" else f"Mpix·samples/s ({workload})",
        "value": round(value, 2),
        "unit": "Mpix·samples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(secs / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "planes": len(lam_d), "out": [d.out_w, d.out_h],
                   "n_samples": d.n_samples, "launch": cfg},
        "roofline": roofline,
        "cpu_baseline": cpu_b,
    }
    if valu is not None:
        line["roofline_valu"] = valu
    if parity is not None:
        line["parity_spot_check"] = parity
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="pixel-c2", choices=sorted(WORKLOADS) + ["batch-c4", "adaptive-c5"])
    ap.add_argument("--no-adaptive", action="store_true",
                    help="skip the learned-budget comparison appended to the default line")
    ap.add_argument("--no-frame", action="store_true",
                    help="skip the host-memory frame comparison appended to the default line")
    ap.add_argument("--no-grain", action="store_true",
                    help="skip the grain-wise C3 summary appended to the default line")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the uneven-batch C4 summary appended to the default line")
    ap.add_argument("--batch-images", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-tuning", action="store_true", help="empty the library's tuned launch table (built-in choices)")
    ap.add_argument("--pix-variant", type=int, default=None,
                    help="pixel-wise formulation: 1 gather, 2 tile splat, 3 HBM splat")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB, so that counter "
                         "runs (FETCH_SIZE, WRITE_SIZE) can be scaled by a copy of known size")
    ap.add_argument("--no-stats", action="store_true",
                    help="skip the instrumented pass (the roofline then has no FLOP count)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "batch-c4":
        b = batch_summary(args.steps, args.warmup, args.batch_images, tuning=not args.no_tuning)
        line = {"metric": b["metric"], "value": b["value"], "unit": b["unit"], "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": b["ms_per_step"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": {"workload": b["workload"], "planes": b["planes"], "pixel_samples": b["pixel_samples"]},
                "roofline": None, "cpu_baseline": None, "batch": b}
        print(json.dumps(line), flush=True)
        return
    if args.workload == "adaptive-c5":
        ad = adaptive_summary(args.steps, args.warmup)
        line = {"metric": ad["metric"], "value": ad["adaptive_ms_per_frame"], "unit": "ms/frame", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": ad["adaptive_ms_per_frame"],
                "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic", "config": {"workload": ad["workload"]}, "roofline": None,
                "cpu_baseline": None, "adaptive": ad}
        print(json.dumps(line), flush=True)
        return

    line = measure(args.workload, args.steps, args.warmup, cpu=not args.no_cpu_baseline,
                   cpu_seconds=args.cpu_seconds, stats=not args.no_stats, tuning=not args.no_tuning,
                   pix_variant=args.pix_variant, copy_calibration_mib=args.copy_calibration_mib)
    if args.workload == "pixel-c2":
        # the other configurations, after the headline's timed region
        if not args.no_grain:
            g = measure("grain-c3", max(2, args.steps), 1, cpu=not args.no_cpu_baseline,
                        cpu_seconds=min(args.cpu_seconds, 8.0))
            line["grain_c3"] = {k: g[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "config",
                                                  "roofline", "roofline_valu", "cpu_baseline") if k in g}
        if not args.no_batch:
            line["batch_c4"] = batch_summary(max(2, args.steps), 1, args.batch_images, tuning=not args.no_tuning)
        if not args.no_frame:
            line["frame"] = frame_summary()
        if not args.no_adaptive:
            line["adaptive"] = adaptive_summary()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
