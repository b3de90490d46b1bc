"""Launch-shape autotuner for the film-grain kernels.

Times candidate launch configurations (tile shape, LDS budget, workgroup size,
formulation) of one integrator on the device with HIP events on the launch
stream and keeps the fastest.  Results are keyed by a workload bucket (path,
radius law, zoom, rm/delta, N) and persisted as a JSON table that
`apply_table` loads into a context before rendering.  All candidates produce
identical output (the parity tests check every formulation), so tuning only
changes speed.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

TABLE_PATH = Path(__file__).resolve().parent / "tuning" / "launch_table.json"


def bucket_key(algo: str, desc) -> str:
    """Workload bucket: configurations tuned for one key are reused for all
    planes in it."""
    nb = 1 << max(0, math.ceil(math.log2(max(1, desc.n_samples))))
    ratio = desc.rm / desc.delta if algo == "pixel" else desc.rm * desc.zoom
    # half-octave buckets of the ratio
    rb = round(2 ** (round(math.log2(max(ratio, 1e-3)) * 2) / 2), 3) if ratio > 0 else 0
    return f"{algo}|{'lognorm' if desc.dist_kind else 'const'}|z{desc.zoom:g}|r{rb:g}|N{nb}"


def pixel_space(desc):
    """Candidate pixel-wise launch shapes for the HBM splat formulation: tile
    shape and workgroup size (the kernel's LDS follows from the tile: one
    N-bit set per pixel plus the sample tables)."""
    words = (desc.n_samples + 31) // 32
    out = []
    for tw, th in [(16, 16), (32, 8), (32, 16), (16, 32), (32, 32), (64, 8), (64, 16), (64, 32),
                   (128, 16), (128, 32)]:
        if tw * th * words * 4 > 128 * 1024:
            continue
        for block in (256, 512):
            out.append(dict(pix_variant=3, pix_tile_w=tw, pix_tile_h=th, pix_block=block))
    return out


def grain_space(desc):
    """Candidate grain-wise launch shapes for the sorted formulation: tile
    shape and workgroup size."""
    words = (desc.n_samples + 31) // 32
    out = []
    for tw, th in [(16, 16), (32, 16), (32, 32), (64, 16), (64, 32), (64, 64), (128, 32), (128, 64)]:
        if tw * th * words * 4 > 96 * 1024:
            continue
        for block in (256, 512):
            out.append(dict(grain_variant=2, grain_tile_w=tw, grain_tile_h=th, grain_block=block))
    return out


def time_config(ctx, algo, desc, lam, offs, out, cfg: dict, reps: int = 3, stream=None) -> float:
    """Median milliseconds of one render with launch config `cfg`."""
    import torch
    stream = stream or torch.cuda.current_stream()
    ctx.set_launch_config(**cfg)
    try:
        ctx.render_device(algo, desc, lam, offs, out, stream)  # warm-up (allocations, checks)
        ctx.stream_status(stream)
        times = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            ctx.render_device(algo, desc, lam, offs, out, stream)
            b.record(stream)
            b.synchronize()
            times.append(a.elapsed_time(b))
        ctx.stream_status(stream)
        return sorted(times)[len(times) // 2]
    finally:
        ctx.reset_launch_config()


def tune(ctx, algo, desc, lam, offs, out, space=None, reps: int = 3, log=None):
    """Returns (best config, [(ms, config), ...] sorted)."""
    space = space if space is not None else (pixel_space(desc) if algo == "pixel" else grain_space(desc))
    results = []
    for cfg in space:
        try:
            ms = time_config(ctx, algo, desc, lam, offs, out, cfg, reps)
        except Exception as e:  # a shape the device rejects is skipped
            if log:
                log(f"[tune] {cfg} rejected: {e}")
            continue
        results.append((ms, cfg))
        if log:
            log(f"[tune] {ms:8.3f} ms  {cfg}")
    results.sort(key=lambda r: r[0])
    return (results[0][1] if results else None), results


def load_table(path: Path = TABLE_PATH) -> dict:
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def save_table(table: dict, path: Path = TABLE_PATH) -> None:
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    Path(path).write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")


def apply_table(ctx, algo, desc, table: dict | None = None) -> dict | None:
    """Sets the tuned launch config for this workload bucket, if one is known."""
    table = load_table() if table is None else table
    cfg = table.get(bucket_key(algo, desc))
    if cfg:
        ctx.set_launch_config(**cfg["config"])
        return cfg["config"]
    return None


def main(argv=None):
    """Tunes one synthetic workload on device 0 and records the winner:
        python -m film_grain_amd.autotune --workload pixel-c2 [--save] [--reps 3]
    """
    import argparse
    import sys

    import numpy as np
    import torch

    from film_grain_amd import capi
    from film_grain_amd.workloads import WORKLOADS, prepare

    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="pixel-c2", choices=sorted(WORKLOADS))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--save", action="store_true", help="merge the winner into the launch table")
    ap.add_argument("--table", default=str(TABLE_PATH))
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    prep = prepare(args.workload)
    dev = torch.device("cuda:0")
    ctx = capi.Context(0)
    lam = torch.from_numpy(prep.lams[0]).to(dev)
    offs = torch.from_numpy(np.ascontiguousarray(prep.offsets)).to(dev)
    out = torch.empty((prep.desc.out_h, prep.desc.out_w), dtype=torch.float32, device=dev)
    best, results = tune(ctx, prep.algo, prep.desc, lam, offs, out, reps=args.reps, log=log)
    ctx.close()
    if best is None:
        log("[tune] no candidate ran")
        return 1
    key = bucket_key(prep.algo, prep.desc)
    entry = {"config": best, "ms": round(results[0][0], 3), "candidates": len(results),
             "workload": WORKLOADS[args.workload]["desc"]}
    print(json.dumps({key: entry}))
    if args.save:
        table = load_table(Path(args.table))
        table[key] = entry
        save_table(table, Path(args.table))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
