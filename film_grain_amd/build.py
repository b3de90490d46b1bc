"""Builds libfilmgrain_hip.so (gfx950 kernels + C ABI + C++ host library)
in-tree with hipcc (the CLI is `python -m film_grain_amd`).  Cross-compiles
without a GPU.

    python -m film_grain_amd.build          # or film_grain_amd.build.build()
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libfilmgrain_hip.so"
TUNING_TABLE = PKG / "tuning" / "launch_table.json"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FG_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the reference never fuses a*b+c, and bit-exact parity with
# its f32/f64 arithmetic depends on it.  Packed FP32 codegen is disabled.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# FG_SANITIZE=address builds the host code (never the GPU code) with
# AddressSanitizer; run Python with the sanitizer runtime preloaded
# (libclang_rt.asan-x86_64.so of /opt/rocm/llvm) and ASAN_OPTIONS=detect_leaks=0.
SANITIZE = os.environ.get("FG_SANITIZE", "")
if SANITIZE:
    COMMON = [f for f in COMMON if f != "-O3"] + ["-O1", "-g", "-Xarch_host", f"-fsanitize={SANITIZE}",
                                                 "-Xarch_host", "-fno-omit-frame-pointer"]
LINK_SANITIZE = ["-Xarch_host", f"-fsanitize={SANITIZE}", "-shared-libasan"] if SANITIZE else []

HIP_SOURCES = ["fg_pixelwise.hip", "fg_pixelwise_hbm.hip", "fg_grainwise.hip", "fg_batch.hip", "fg_mlp.hip",
               "fg_frame.hip", "fg_probe.hip", "fg_capi.hip"]
# The exact-logf correction table (fg_logf.h) is generated from this
# machine's C library at every build and checked over its whole domain.
LOGF_GEN = "gen_logf_table.hip"
LOGF_TABLE = "fg_logf_table.h"
HOST_SOURCES = sorted(p.name for p in (CSRC / "host").glob("*.cpp")) if (CSRC / "host").exists() else []


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    if r.stderr.strip():
        # keep real warnings visible; hipcc's host pass notes the unknown
        # packed-fp32 feature, which is expected
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


def logf_table() -> Path:
    """Builds and runs the logf table generator; returns the header."""
    gen = BUILD / "gen_logf_table"
    _run([HIPCC, *COMMON, *DEVICE, "-fno-builtin", str(CSRC / LOGF_GEN), "-o", str(gen), "-lm", "-lpthread"])
    header = BUILD / LOGF_TABLE
    r = subprocess.run([str(gen), str(header)], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("logf table generation failed")
    return header


EXP_GEN = "gen_exp_table.cpp"
EXP_TABLE = "fg_exp_neg_table.h"
GXX = os.environ.get("CXX", "g++")


def exp_table() -> Path:
    """Builds and runs the exp(-x) table generator (host C++ only); returns
    the header."""
    gen = BUILD / "gen_exp_table"
    _run([GXX, "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-pthread", f"-I{CSRC}",
          str(CSRC / EXP_GEN), "-o", str(gen), "-lm"])
    header = BUILD / EXP_TABLE
    r = subprocess.run([str(gen), str(header)], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("exp table generation failed")
    return header


LAUNCH_FIELDS = ["pix_tile_w", "pix_tile_h", "pix_lds_bytes", "grain_tile_w", "grain_tile_h", "grain_lds_bytes",
                 "grain_grains_per_thread", "pix_variant", "pix_block", "grain_variant", "grain_block"]


def parse_bucket(key: str) -> tuple[int, int, int, int, int]:
    """'pixel|lognorm|zq4|rh4|N256' -> (algo, dist_kind, zoom_q, ratio_h, n_bucket)."""
    algo, dist, zq, rh, nb = key.split("|")
    if algo not in ("pixel", "grain") or dist not in ("const", "lognorm") or not zq.startswith("zq") \
            or not rh.startswith("rh") or not nb.startswith("N"):
        raise ValueError(f"bad launch table key {key!r}")
    return (0 if algo == "pixel" else 1, 0 if dist == "const" else 1, int(zq[2:]), int(rh[2:]), int(nb[1:]))


def launch_table_header() -> Path:
    """The tuned launch table as the library's built-in table (C header)."""
    import json
    table = json.loads(TUNING_TABLE.read_text()) if TUNING_TABLE.exists() else {}
    rows = []
    for key in sorted(table):
        algo, dist, zq, rh, nb = parse_bucket(key)
        cfg = table[key]["config"]
        unknown = set(cfg) - set(LAUNCH_FIELDS)
        if unknown:
            raise ValueError(f"launch table entry {key!r}: unknown fields {sorted(unknown)}")
        vals = ", ".join(str(int(cfg.get(f, 0))) for f in LAUNCH_FIELDS)
        rows.append(f"    {{{{{algo}u, {dist}u, {zq}, {rh}, {nb}u}}, {{{vals}}}}},  // {key}")
    header = BUILD / "fg_launch_table.h"
    text = ("// Generated by film_grain_amd/build.py from tuning/launch_table.json: the\n"
            "// library's built-in tuned launch table (fg_launch_entry, declared in\n"
            "// film_grain_hip.h, which the including file includes first).\n"
            "#pragma once\n"
            "static const fg_launch_entry kBuiltinLaunchTable[] = {\n"
            + ("\n".join(rows) if rows else "    {}") + "\n};\n"
            f"static const int kBuiltinLaunchCount = {len(rows)};\n")
    if not header.exists() or header.read_text() != text:
        header.write_text(text)
    return header


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    logf_table()
    exp_table()
    launch_table_header()
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        jobs.append([HIPCC, *COMMON, *DEVICE, f"-I{BUILD}", "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = BUILD / ("host_" + src + ".o")
        objs.append(obj)
        jobs.append([HIPCC, *COMMON, f"-I{BUILD}", "-c", str(CSRC / "host" / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, *DEVICE, "-shared", "-fPIC", *LINK_SANITIZE, "-o", str(tmp), *map(str, objs), "-lz", "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    for p in (LIB,):
        if p.exists():
            p.unlink()


if __name__ == "__main__":
    build(verbose=True)
