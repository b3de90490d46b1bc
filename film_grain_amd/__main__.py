"""Command line of the MI355X backend, with the reference's flags
(src/main.rs:71-247) and output lines (src/main.rs:249-275):

    python -m film_grain_amd INPUT -o OUTPUT [--radius R] [--radius-dist const|lognorm]
        [--radius-stddev SD] [--zoom S] [--sigma PX] [--iters N] [--color-mode rgb|luma]
        [--device hip|gpu|cpu] [--algo auto|grain|pixel] [--max-radius RM] [--cell DELTA]
        [--roi X0,Y0:X1,Y1] [--size WxH] [--format EXT] [--dry-run] [--explain] [--seed SEED]
        [--route reference|measured]

--device defaults to hip here (the reference's default is cpu, whose
integrators this backend does not include).  --explain adds, on hip, each
integrator's launch shape and its time per frame, estimated from a timed
render of a centred crop of the frame.  --route measured resolves --algo
auto to the integrator so measured faster instead of the reference's
heuristic (src/choose.rs:4-26);
the two integrators draw different random streams, so this changes the
image, which is why it is not the default.
"""
from __future__ import annotations

import argparse
import sys


def _parser():
    ap = argparse.ArgumentParser(prog="python -m film_grain_amd",
                                 description="Physically based film grain on MI355X (hip).")
    ap.add_argument("input", metavar="INPUT")
    ap.add_argument("-o", "--output", metavar="OUTPUT", required=True)
    ap.add_argument("--radius", type=float, default=0.10, metavar="R")
    ap.add_argument("--radius-dist", choices=["const", "lognorm"], default="const", metavar="KIND")
    ap.add_argument("--radius-stddev", type=float, default=0.0, metavar="SD")
    ap.add_argument("--zoom", type=float, default=1.0, metavar="S")
    ap.add_argument("--sigma", type=float, default=0.8, metavar="PX")
    ap.add_argument("--iters", type=int, default=32, metavar="N")
    ap.add_argument("--color-mode", choices=["rgb", "luma"], default="luma", metavar="M")
    ap.add_argument("--device", choices=["hip", "gpu", "cpu"], default="hip", metavar="DEVICE")
    ap.add_argument("--algo", choices=["auto", "grain", "pixel"], default="auto", metavar="A")
    ap.add_argument("--max-radius", default="quantile 0.999", metavar="RM")
    ap.add_argument("--cell", default="approx mu_r", metavar="DELTA")
    ap.add_argument("--roi", default=None, metavar="X0,Y0:X1,Y1")
    ap.add_argument("--size", default=None, metavar="WxH")
    ap.add_argument("--format", default=None, metavar="EXT")
    ap.add_argument("--dry-run", action="store_true")
    ap.add_argument("--explain", action="store_true")
    ap.add_argument("--seed", type=int, default=5489, metavar="SEED")
    ap.add_argument("--route", choices=["reference", "measured"], default="reference")
    return ap


def _algo_name(code: int) -> str:
    return {1: "Grain", 2: "Pixel"}.get(int(code), "Auto")


def main(argv=None) -> int:
    a = _parser().parse_args(argv)
    from film_grain_amd import host, route
    if a.radius_dist == "const" and a.radius_stddev > 0.0:
        print("warning: --radius-sd is ignored when --radius-dist=const", file=sys.stderr)
    args = host.CliArgs(input_path=a.input, output_path=a.output, radius_dist=a.radius_dist, radius=a.radius,
                        radius_stddev=a.radius_stddev, zoom=a.zoom, sigma=a.sigma, iters=a.iters, algo=a.algo,
                        max_radius=a.max_radius, cell=a.cell, color_mode=a.color_mode, device=a.device,
                        roi=a.roi, size=a.size, seed=a.seed, dry_run=a.dry_run, explain=a.explain,
                        format=a.format)
    try:
        info = None
        if a.explain or (a.route == "measured" and a.algo == "auto"):
            info = route.explain(args, host.load_image(a.input))
            if a.route == "measured" and a.algo == "auto" and "measured_choice" in info:
                args.algo = info["measured_choice"]
        stats = host.dry_run(args) if a.dry_run else host.render_file(args)
    except host.RenderError as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    if a.dry_run:
        print(f"dry-run: {_algo_name(stats.algorithm)} algorithm, input {stats.in_w}x{stats.in_h} → output "
              f"{stats.out_w}x{stats.out_h}, samples {stats.n_samples}")
    else:
        print(f"rendered {_algo_name(stats.algorithm)} algorithm → {a.output} ({stats.out_w}x{stats.out_h})")
    if a.explain:
        dev = {"cpu": "Cpu", "gpu": "Hip", "hip": "Hip"}[a.device]
        print(f"sigma/mean {stats.sigma_ratio:.3f}, rm/mean {stats.rm_ratio:.3f}, samples {stats.n_samples} "
              f"on {dev}")
        if info and "hip" in info:
            parts = []
            for algo in ("pixel", "grain"):
                h = info["hip"][algo]
                launch = ",".join(f"{k}={v}" for k, v in h["launch"].items())
                parts.append(f"{algo}-wise [{launch}] ~{h['estimated_ms_per_frame']:.2f} ms/frame")
            print("hip: " + "; ".join(parts) + f" (crop {info['probe_crop']} measured); faster: "
                  f"{info['measured_choice']}; route {a.route}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
