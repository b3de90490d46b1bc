"""Algorithm routing and `--explain` for the hip device.

The reference routes `--algo auto` with a parameter heuristic
(src/choose.rs:4-26) and `--explain` prints the ratios behind it
(src/lib.rs:205-222, src/main.rs:249-275).  That choice is part of the
output: the two integrators draw different random streams, so they render
different (equally valid) images, and the default keeps the reference's
choice.  For the hip device this module adds what the reference's explain
help promises ("incl. estimated cost"), measured rather than modelled: each
integrator renders a centred crop of the actual frame (at most 1024 x 1024
input pixels, its first rendered plane, the actual parameters) on the
device, and the crop's time, scaled by the frame's area over the crop's and
by the planes rendered, estimates the frame.  A linear model of both kernels
in counted work terms, fitted to a grid of planes, missed the grain-wise
times by 39% at the median and by 2x on the pixel-wise benchmark frame; the
crop measurement sees the actual content and launch shapes.  The opt-in
router (`--route measured`) picks the faster integrator.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

PROBE_SIDE = 1024


def crop_box(w: int, h: int, side: int = PROBE_SIDE):
    """Centred crop of at most side x side: (x0, y0, x1, y1)."""
    cw, ch = min(w, side), min(h, side)
    x0, y0 = (w - cw) // 2, (h - ch) // 2
    return x0, y0, x0 + cw, y0 + ch


def roi_crop(args, rgb):
    """The frame after the --roi crop (src/color.rs:215-231)."""
    if not args.roi:
        return rgb
    (x0, y0), (x1, y1) = ([int(v) for v in part.split(",")] for part in args.roi.split(":"))
    return rgb[y0:y1, x0:x1]


def probe(ctx, args, rgb: np.ndarray, reps: int = 2) -> dict:
    """Estimated ms per frame of each integrator on the device, from a timed
    render of a centred crop of the frame's first rendered plane."""
    import torch
    from film_grain_amd import capi, host
    rgb = np.ascontiguousarray(roi_crop(args, rgb), dtype=np.float32)
    h, w = rgb.shape[:2]
    planes = host.color_split(rgb, args.color_mode)
    n_planes = 3 if args.color_mode == "rgb" else 1
    x0, y0, x1, y1 = crop_box(w, h)
    u = np.ascontiguousarray(planes[0][y0:y1, x0:x1])
    a = dataclasses.replace(args, roi=None, size=None)
    d, offs, offs_in = host.derive(a, x1 - x0, y1 - y0)
    nrm, _ = host.normalize_plane(u)
    lam = torch.from_numpy(host.lambda_plane(nrm, d.inv_e_pi_r2)).to("cuda:0")
    desc = capi.make_desc(d.in_w, d.in_h, d.out_w, d.out_h, d.n_samples, d.dist_kind, d.seed, d.zoom, d.delta,
                          d.rm, d.radius_mean, d.log_mu, d.log_sigma)
    out = torch.empty((d.out_h, d.out_w), dtype=torch.float32, device="cuda:0")
    scale = (w * h) / float((x1 - x0) * (y1 - y0)) * n_planes
    est = {}
    for algo, o in (("pixel", offs_in), ("grain", offs)):
        od = torch.from_numpy(np.ascontiguousarray(o)).to("cuda:0")
        ctx.render_device(algo, desc, lam, od, out)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.render_device(algo, desc, lam, od, out)
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        est[algo] = float(np.median(ts)) * scale
    ctx.stream_status()
    return {"crop": [x0, y0, x1, y1], "ms_per_frame": est}


def explain(args, rgb: np.ndarray) -> dict:
    """The reference's routing facts for this render plus, on the hip device,
    per integrator: the launch shape it would run on the frame and the
    crop-measured estimate of its time; 'measured_choice' is the faster."""
    from film_grain_amd import capi, host
    frame = roi_crop(args, rgb)
    h, w = frame.shape[:2]
    d, _, _ = host.derive(dataclasses.replace(args, roi=None), w, h)
    out = {"algorithm": {1: "grain", 2: "pixel"}.get(int(d.algorithm), "?"), "sigma_ratio": float(d.sigma_ratio),
           "rm_ratio": float(d.rm_ratio), "n_samples": int(d.n_samples), "device": args.device}
    if args.device not in ("hip", "gpu"):
        return out
    desc = capi.make_desc(d.in_w, d.in_h, d.out_w, d.out_h, d.n_samples, d.dist_kind, d.seed, d.zoom, d.delta,
                          d.rm, d.radius_mean, d.log_mu, d.log_sigma)
    ctx = capi.Context(0)
    try:
        hip = {}
        for algo in ("pixel", "grain"):
            cfg = ctx.launch_config(desc, algo)
            hip[algo] = {"launch": {k: v for k, v in cfg.items()
                                    if v and k.startswith("pix" if algo == "pixel" else "grain")}}
        pr = probe(ctx, args, rgb)
    finally:
        ctx.close()
    for algo in ("pixel", "grain"):
        hip[algo]["estimated_ms_per_frame"] = round(pr["ms_per_frame"][algo], 3)
    out["hip"] = hip
    out["probe_crop"] = pr["crop"]
    out["measured_choice"] = min(("pixel", "grain"), key=lambda a: pr["ms_per_frame"][a])
    return out


def main(argv=None):
    """python -m film_grain_amd.route check: crop estimates against full
    renders of three planes (on the device); one JSON line per plane."""
    import argparse
    import json

    import torch
    from film_grain_amd import capi, host, synth
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["check"])
    ap.parse_args(argv)
    ctx = capi.Context(0)
    for name, u, a in (
            ("1920x1080 zoom 2, lognormal 0.1/0.05, N=256", synth.c2_rgb(7, 1920, 1080)[0],
             host.CliArgs(radius=0.1, radius_stddev=0.05, radius_dist="lognorm", zoom=2.0, iters=256, seed=7)),
            ("4096x4096 ramp, r=0.05, N=128", synth.c3_ramp(11, 4096),
             host.CliArgs(radius=0.05, zoom=1.0, iters=128, seed=11)),
            ("2048x2048 sinusoid, r=0.25, N=64, zoom 1", synth.sinusoid_field(2048, 2048, 5, n_waves=4, noise=0.02),
             host.CliArgs(radius=0.25, zoom=1.0, iters=64, seed=5))):
        a.color_mode = "luma"
        rgb = np.ascontiguousarray(np.repeat(u[:, :, None], 3, axis=2))
        est = probe(ctx, a, rgb)["ms_per_frame"]
        d, offs, offs_in = host.derive(a, u.shape[1], u.shape[0])
        y = host.color_split(rgb, "luma")[0]
        nrm, _ = host.normalize_plane(y)
        lam = torch.from_numpy(host.lambda_plane(nrm, d.inv_e_pi_r2)).to("cuda:0")
        desc = capi.make_desc(d.in_w, d.in_h, d.out_w, d.out_h, d.n_samples, d.dist_kind, d.seed, d.zoom, d.delta,
                              d.rm, d.radius_mean, d.log_mu, d.log_sigma)
        out = torch.empty((d.out_h, d.out_w), dtype=torch.float32, device="cuda:0")
        meas = {}
        for algo, o in (("pixel", offs_in), ("grain", offs)):
            od = torch.from_numpy(np.ascontiguousarray(o)).to("cuda:0")
            ctx.render_device(algo, desc, lam, od, out)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.render_device(algo, desc, lam, od, out)
            torch.cuda.synchronize()
            meas[algo] = (time.perf_counter() - t0) * 1e3
        print(json.dumps({"plane": name, "estimated_ms": est, "measured_ms": meas,
                          "choice_estimated": min(est, key=est.get), "choice_measured": min(meas, key=meas.get)}),
              flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
