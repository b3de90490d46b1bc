"""Builds libfilmgrain_hip.so (gfx950 kernels + C ABI + C++ host library) and
the filmgrain CLI, in-tree, with hipcc.  Cross-compiles without a GPU.

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
CLI = PKG / "bin" / "filmgrain"

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


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    logf_table()
    exp_table()
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
    cli_src = CSRC / "cli" / "filmgrain_main.cpp"
    if cli_src.exists():
        CLI.parent.mkdir(exist_ok=True)
        _run([HIPCC, *COMMON, "-o", str(CLI), str(cli_src), f"-L{PKG}", "-lfilmgrain_hip",
              "-Wl,-rpath,$ORIGIN/..", "-lz", "-lpthread"])
    if verbose:
        print(f"built {LIB}")
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    for p in (LIB, CLI):
        if p.exists():
            p.unlink()


if __name__ == "__main__":
    build(verbose=True)
