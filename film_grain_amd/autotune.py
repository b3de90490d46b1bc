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
    """Workload bucket of a plane, as the library computes it
    (fg_launch_bucket): configurations tuned for one key serve every plane
    in it."""
    from film_grain_amd import capi
    return capi.bucket_name(capi.launch_bucket(desc, algo))


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


def table_entries(table: dict) -> list:
    """The JSON table as fg_set_launch_table entries: (LaunchKey, config)."""
    from film_grain_amd import capi
    from film_grain_amd.build import parse_bucket
    out = []
    for key, entry in sorted(table.items()):
        algo, dist, zq, rh, nb = parse_bucket(key)
        out.append((capi.LaunchKey(algo, dist, zq, rh, nb), dict(entry["config"])))
    return out


def apply_table(ctx, table: dict | None = None) -> int:
    """Installs a table (default: the JSON file as it is now, which may be
    newer than the table the library was built with) into a context; returns
    its size.  The library's own table serves every context without this."""
    table = load_table() if table is None else table
    entries = table_entries(table)
    ctx.set_launch_table(entries)
    return len(entries)


def tune_batch_c4(ctx, reps: int, log) -> dict:
    """One entry per bucket of the uneven batch's planes (64 images, sides
    128-2048, mixed radii, N and channels): each bucket tuned on its largest
    plane."""
    import numpy as np
    import torch
    from film_grain_amd import workloads
    planes = workloads.prepare_batch(workloads.batch_images(master_seed=2024, n_images=64))
    best_plane = {}
    for d, lam, offs in planes:
        key = bucket_key("pixel", d)
        if key not in best_plane or d.out_w * d.out_h > best_plane[key][0].out_w * best_plane[key][0].out_h:
            best_plane[key] = (d, lam, offs)
    dev = torch.device("cuda:0")
    out_table = {}
    for key, (d, lam, offs) in sorted(best_plane.items()):
        lam_t = torch.from_numpy(lam).to(dev)
        offs_t = torch.from_numpy(np.ascontiguousarray(offs)).to(dev)
        out = torch.empty((d.out_h, d.out_w), dtype=torch.float32, device=dev)
        best, results = tune(ctx, "pixel", d, lam_t, offs_t, out, reps=reps)
        if best is None:
            continue
        out_table[key] = {"config": best, "ms": round(results[0][0], 3), "candidates": len(results),
                          "workload": f"uneven batch plane {d.in_w}x{d.in_h}, N={d.n_samples}"}
        log(f"[tune] {key}: {best} {results[0][0]:.3f} ms (of {len(results)})")
    return out_table


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
    ap.add_argument("--workload", default="pixel-c2", choices=sorted(WORKLOADS) + ["batch-c4"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--save", action="store_true", help="merge the winner into the launch table")
    ap.add_argument("--table", default=str(TABLE_PATH))
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.workload == "batch-c4":
        ctx = capi.Context(0)
        ctx.set_launch_table([])  # time the candidates, not the table
        found = tune_batch_c4(ctx, args.reps, log)
        ctx.close()
        print(json.dumps(found))
        if args.save:
            table = load_table(Path(args.table))
            table.update(found)
            save_table(table, Path(args.table))
        return 0
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
