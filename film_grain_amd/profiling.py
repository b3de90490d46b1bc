"""Summaries of rocprofv3 output for the kernels of this library.

rocprofv3 writes its results as SQLite databases (`<dir>/<name>_results.db`).
This module reads a kernel trace and, optionally, a FETCH_SIZE and a
WRITE_SIZE counter run of the same command, and reports per kernel the average
duration and the HBM bytes per dispatch.  FETCH_SIZE and WRITE_SIZE (KiB per
dispatch) are scaled by what the same counters read for a device-to-device
copy of known size that the profiled command performs first (bench.py
--copy-calibration-mib M): factor = counter KiB / (M * 1024).  Further counter
runs (--pmc, any counters) add a table of each counter's average per
dispatch for the longest kernels.

    python -m film_grain_amd.profiling --trace T [--fetch F --write W --copy-mib 256] \
        --kernel splat_hbm_kernel [--pmc P ...] --markdown out.md [--json traffic.json --key K]
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import sqlite3
import statistics


def _db(path: str) -> str:
    if os.path.isfile(path):
        return path
    hits = sorted(glob.glob(os.path.join(path, "**", "*_results.db"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no rocprofv3 results database under {path}")
    return hits[0]


def kernel_times(trace: str) -> dict:
    """kernel name -> (calls, average microseconds) from a --kernel-trace
    database (its top_kernels view reports microseconds)."""
    con = sqlite3.connect(_db(trace))
    out = {}
    for name, calls, avg in con.execute("select name, total_calls, average from top_kernels"):
        out[name] = (int(calls), float(avg))
    return out


def counter_values(db: str, counter: str) -> dict:
    """kernel name -> list of per-dispatch values of one counter."""
    con = sqlite3.connect(_db(db))
    out: dict = {}
    for name, v in con.execute("select kernel_name, value from counters_collection where counter_name = ? "
                               "order by dispatch_id", (counter,)):
        out.setdefault(name, []).append(float(v))
    return out


def copy_factor(values: dict, copy_kib: float) -> float:
    """Counter reading of the calibration copy over its size: the largest
    dispatch of the runtime's copy kernel."""
    best = 0.0
    for name, vals in values.items():
        if "copyBuffer" in name:
            best = max(best, max(vals))
    if best <= 0.0:
        raise ValueError("no calibration copy found in the counter run")
    return best / copy_kib


def counter_table(dbs: list[str]) -> dict:
    """kernel name -> {counter: average per dispatch} over counter runs."""
    out: dict = {}
    for db in dbs:
        con = sqlite3.connect(_db(db))
        vals: dict = {}
        for name, cn, v in con.execute("select kernel_name, counter_name, value from counters_collection"):
            vals.setdefault((name, cn), []).append(float(v))
        for (name, cn), v in vals.items():
            out.setdefault(name, {})[cn] = statistics.fmean(v)
    return out


def summarize(trace: str, fetch: str | None, write: str | None, copy_mib: int, kernel: str,
              pmc: list[str] | None = None) -> dict:
    times = kernel_times(trace)
    fv = counter_values(fetch, "FETCH_SIZE") if fetch else {}
    wv = counter_values(write, "WRITE_SIZE") if write else {}
    kib = copy_mib * 1024.0
    ff = copy_factor(fv, kib) if fetch else None
    wf = copy_factor(wv, kib) if write else None
    rows = []
    for name, (calls, avg_us) in sorted(times.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        f = statistics.fmean(fv[name]) if name in fv else None
        w = statistics.fmean(wv[name]) if name in wv else None
        traffic = None
        if f is not None and w is not None:
            traffic = (f / ff + w / wf) * 1024.0
        rows.append(dict(kernel=name, calls=calls, avg_us=avg_us, fetch_kib=f, write_kib=w,
                         traffic_bytes=traffic))
    main = [r for r in rows if kernel in r["kernel"]]
    return dict(fetch_factor=ff, write_factor=wf, copy_mib=copy_mib, rows=rows,
                kernel=main[0] if main else None, counters=counter_table(pmc or []))


def markdown(s: dict, title: str, commands: list[str], notes: list[str]) -> str:
    lines = [f"# {title}", "", "Commands (one MI355X; `R` is the repository root, `S` any scratch directory):", "",
             "```"]
    lines += commands
    lines += ["```", ""]
    if s["fetch_factor"] is not None and s["write_factor"] is not None:
        lines += [f"Counter calibration: a {s['copy_mib']} MiB device-to-device copy reads "
                  f"FETCH_SIZE = {s['fetch_factor']:.4f} x its bytes and WRITE_SIZE = "
                  f"{s['write_factor']:.4f} x its bytes; per-kernel traffic below divides by these factors.", ""]
    lines += ["| kernel | calls | avg us | FETCH KiB/dispatch | WRITE KiB/dispatch | HBM bytes/dispatch (scaled) |",
              "|---|---:|---:|---:|---:|---:|"]
    for r in s["rows"]:
        if r["avg_us"] < 1.0 and r["fetch_kib"] is None:
            continue
        f = "" if r["fetch_kib"] is None else f"{r['fetch_kib']:.0f}"
        w = "" if r["write_kib"] is None else f"{r['write_kib']:.0f}"
        t = "" if r["traffic_bytes"] is None else f"{r['traffic_bytes']:.4g}"
        lines.append(f"| `{r['kernel'][:90]}` | {r['calls']} | {r['avg_us']:.1f} | {f} | {w} | {t} |")
    counters = s.get("counters") or {}
    if counters:
        names = sorted({c for d in counters.values() for c in d})
        top = [r["kernel"] for r in s["rows"] if r["kernel"] in counters][:6]
        lines += ["", "Counters, average per dispatch (SQ_BUSY_CYCLES is summed over the shader engines):", "",
                  "| kernel | " + " | ".join(names) + " |", "|---|" + "---:|" * len(names)]
        for k in top:
            lines.append(f"| `{k[:60]}` | " + " | ".join(f"{counters[k].get(n, float('nan')):.4g}" for n in names)
                         + " |")
    lines += [""] + notes + [""]
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--pmc", action="append", default=[], help="a counter run's output (repeatable)")
    ap.add_argument("--kernel", required=True, help="substring of the dominant kernel's name")
    ap.add_argument("--markdown")
    ap.add_argument("--title", default="Kernel profile")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--note", action="append", default=[])
    ap.add_argument("--json", help="merge {key: traffic record} into this JSON file")
    ap.add_argument("--key")
    args = ap.parse_args(argv)
    s = summarize(args.trace, args.fetch, args.write, args.copy_mib, args.kernel, args.pmc)
    if args.markdown:
        with open(args.markdown, "w") as f:
            f.write(markdown(s, args.title, args.command, args.note))
    k = s["kernel"]
    print(json.dumps(dict(kernel=k, fetch_factor=s["fetch_factor"], write_factor=s["write_factor"])))
    if args.json and args.key and k and k["traffic_bytes"] is not None:
        try:
            table = json.load(open(args.json))
        except (OSError, ValueError):
            table = {}
        table[args.key] = dict(kernel=k["kernel"], avg_us=round(k["avg_us"], 2),
                               traffic_bytes=round(k["traffic_bytes"]), fetch_factor=round(s["fetch_factor"], 5),
                               write_factor=round(s["write_factor"], 5))
        with open(args.json, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
            f.write("\n")


if __name__ == "__main__":
    main()
