"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle restates the reference's CPU integrators (src/pixelwise.rs,
src/grainwise.rs, src/rng.rs, src/model.rs) in plain C; see fg_oracle.c for the
per-function reference citations and pinning status.  Only tests/, the smoke
check and the benchmark's CPU-baseline leg import this module, and only to check
or time the CPU path; the product under film_grain_amd/ never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libfg_oracle.so"


class FgoParams(C.Structure):
    _fields_ = [
        ("radius_mean", C.c_float),
        ("radius_stddev", C.c_float),
        ("dist_kind", C.c_uint32),
        ("zoom", C.c_float),
        ("sigma_px", C.c_float),
        ("n_samples", C.c_uint32),
        ("max_radius_is_abs", C.c_int32),
        ("max_radius_value", C.c_float),
        ("cell_delta", C.c_float),
        ("seed", C.c_uint64),
        ("in_w", C.c_uint32),
        ("in_h", C.c_uint32),
        ("size_w", C.c_int32),
        ("size_h", C.c_int32),
    ]


class FgoDerived(C.Structure):
    _fields_ = [
        ("in_w", C.c_uint32),
        ("in_h", C.c_uint32),
        ("out_w", C.c_uint32),
        ("out_h", C.c_uint32),
        ("inv_e_pi_r2", C.c_float),
        ("rm", C.c_float),
        ("delta", C.c_float),
        ("radius_stddev", C.c_float),
        ("log_mu", C.c_float),
        ("log_sigma", C.c_float),
    ]


class FgoPlane(C.Structure):
    _fields_ = [
        ("in_w", C.c_uint32),
        ("in_h", C.c_uint32),
        ("out_w", C.c_uint32),
        ("out_h", C.c_uint32),
        ("n_samples", C.c_uint32),
        ("dist_kind", C.c_uint32),
        ("seed", C.c_uint64),
        ("zoom", C.c_float),
        ("delta", C.c_float),
        ("rm", C.c_float),
        ("radius_mean", C.c_float),
        ("log_mu", C.c_float),
        ("log_sigma", C.c_float),
    ]


class FgoCounters(C.Structure):
    _fields_ = [
        ("points", C.c_uint64),
        ("cell_visits", C.c_uint64),
        ("tests", C.c_uint64),
        ("tests_full", C.c_uint64),
        ("hits", C.c_uint64),
    ]


class FgoGrainCounters(C.Structure):
    _fields_ = [("grains", C.c_uint64), ("pairs", C.c_uint64), ("tests", C.c_uint64), ("hits", C.c_uint64)]


_lib = None
_U64P = C.POINTER(C.c_uint64)
_F32P = C.POINTER(C.c_float)
_F64P = C.POINTER(C.c_double)
_U32P = C.POINTER(C.c_uint32)


def build() -> None:
    """Compile the oracle with its Makefile (gcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    L.fgo_splitmix64.argtypes = [C.c_uint64]
    L.fgo_splitmix64.restype = C.c_uint64
    L.fgo_mix.argtypes = [C.c_uint64, C.c_uint64]
    L.fgo_mix.restype = C.c_uint64
    L.fgo_mix3.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_int64]
    L.fgo_mix3.restype = C.c_uint64
    L.fgo_xoshiro_seed.argtypes = [C.c_uint64, _U64P]
    L.fgo_xoshiro_next_u64.argtypes = [_U64P]
    L.fgo_xoshiro_next_u64.restype = C.c_uint64
    L.fgo_xoshiro_next_u32.argtypes = [_U64P]
    L.fgo_xoshiro_next_u32.restype = C.c_uint32
    L.fgo_uniform_f32_scale.argtypes = [C.c_float, C.c_float]
    L.fgo_uniform_f32_scale.restype = C.c_float
    L.fgo_uniform_f32_sample.argtypes = [_U64P, C.c_float, C.c_float]
    L.fgo_uniform_f32_sample.restype = C.c_float
    L.fgo_standard_f64.argtypes = [_U64P]
    L.fgo_standard_f64.restype = C.c_double
    L.fgo_open01_f64.argtypes = [_U64P]
    L.fgo_open01_f64.restype = C.c_double
    L.fgo_log_gamma.argtypes = [C.c_double]
    L.fgo_log_gamma.restype = C.c_double
    L.fgo_poisson_sample.argtypes = [_U64P, C.c_double]
    L.fgo_poisson_sample.restype = C.c_double
    L.fgo_standard_normal.argtypes = [_U64P]
    L.fgo_standard_normal.restype = C.c_double
    L.fgo_lognormal_sample.argtypes = [_U64P, C.c_double, C.c_double]
    L.fgo_lognormal_sample.restype = C.c_double
    L.fgo_zig_tables.argtypes = [_F64P, _F64P]
    L.fgo_chacha_block.argtypes = [_U32P, C.c_uint64, C.c_uint64, C.c_int, _U32P]
    L.fgo_pcg32_next.argtypes = [_U64P, C.c_uint64]
    L.fgo_pcg32_next.restype = C.c_uint32
    L.fgo_stdrng_u64.argtypes = [C.c_uint64, C.c_uint32, _U64P]
    L.fgo_make_offsets.argtypes = [C.c_uint64, C.c_uint32, C.c_float, _F32P]
    L.fgo_make_offsets.restype = C.c_int
    L.fgo_inverse_normal_cdf.argtypes = [C.c_double]
    L.fgo_inverse_normal_cdf.restype = C.c_double
    L.fgo_derive.argtypes = [C.POINTER(FgoParams), C.POINTER(FgoDerived)]
    L.fgo_derive.restype = C.c_int
    L.fgo_default_cell_delta.argtypes = [C.c_float]
    L.fgo_default_cell_delta.restype = C.c_float
    L.fgo_choose_algorithm.argtypes = [C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_int]
    L.fgo_choose_algorithm.restype = C.c_int
    L.fgo_normalize_plane.argtypes = [_F32P, C.c_uint64, _F32P]
    L.fgo_normalize_plane.restype = C.c_float
    L.fgo_lambda_plane.argtypes = [_F32P, C.c_uint64, C.c_float, _F32P]
    L.fgo_render_pixelwise.argtypes = [C.POINTER(FgoPlane), _F32P, _F32P, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_uint32, _F32P, C.c_int,
                                       C.POINTER(FgoCounters)]
    L.fgo_render_pixelwise.restype = C.c_int
    L.fgo_render_grainwise.argtypes = [C.POINTER(FgoPlane), _F32P, _F32P, _F32P, _U64P, C.c_int,
                                       _U64P]
    L.fgo_render_grainwise.restype = C.c_int
    L.fgo_render_grainwise_counted.argtypes = [C.POINTER(FgoPlane), _F32P, _F32P, _F32P, _U64P, C.c_int,
                                               C.POINTER(FgoGrainCounters)]
    L.fgo_render_grainwise_counted.restype = C.c_int
    L.fgo_cell_dump.argtypes = [C.POINTER(FgoPlane), _F32P, C.c_int32, C.c_int32, _U64P, _U64P,
                                _F32P, C.c_int]
    L.fgo_cell_dump.restype = C.c_int
    L.fgo_cell_count.argtypes = [C.POINTER(FgoPlane), _F32P, C.c_int32, C.c_int32]
    L.fgo_cell_count.restype = C.c_uint32
    L.fgo_pixel_dump.argtypes = [C.POINTER(FgoPlane), _F32P, C.c_int32, C.c_int32, _U64P, _U64P,
                                 _F32P, C.c_int]
    L.fgo_pixel_dump.restype = C.c_int
    L.fgo_max_threads.restype = C.c_int
    L.fgo_cell_digests.argtypes = [C.POINTER(FgoPlane), _F32P, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32,
                                   _U32P, _U64P, C.c_int]
    L.fgo_cell_digests.restype = C.c_int
    L.fgo_pixel_counts.argtypes = [C.POINTER(FgoPlane), _F32P, _U32P, _U64P, C.c_int]
    L.fgo_pixel_counts.restype = C.c_int
    L.fgo_glibc_math.argtypes = [C.c_int, _F64P, C.c_int64, _F64P, C.c_int]
    L.fgo_glibc_math.restype = C.c_int
    L.fgo_to_u8.argtypes = [C.c_float]
    L.fgo_to_u8.restype = C.c_uint8
    L.fgo_color_split.argtypes = [C.c_int, _F32P, C.c_uint64, _F32P, _F32P, _F32P]
    L.fgo_color_split.restype = C.c_int
    L.fgo_resize_nearest.argtypes = [_F32P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _F32P]
    L.fgo_resize_nearest.restype = None
    L.fgo_color_merge.argtypes = [C.c_int, _F32P, _F32P, _F32P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.POINTER(C.c_uint8)]
    L.fgo_color_merge.restype = C.c_int
    L.fgo_color_constants.argtypes = [_F32P]
    L.fgo_color_constants.restype = None
    _lib = L
    return L


def _f32p(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F32P)


def _u64p(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(_U64P)


# ---- thin pythonic helpers -------------------------------------------------

class Xoshiro:
    """Xoshiro256++ with rand 0.8.5 seed_from_u64 (oracle)."""

    def __init__(self, seed: int | None = None, state=None):
        self.s = np.zeros(4, dtype=np.uint64)
        if state is not None:
            self.s[:] = np.array(state, dtype=np.uint64)
        else:
            lib().fgo_xoshiro_seed(C.c_uint64(seed), _u64p(self.s))

    def next_u64(self) -> int:
        return int(lib().fgo_xoshiro_next_u64(_u64p(self.s)))

    def next_u32(self) -> int:
        return int(lib().fgo_xoshiro_next_u32(_u64p(self.s)))

    def uniform_f32(self, low: float, scale: float) -> float:
        return lib().fgo_uniform_f32_sample(_u64p(self.s), low, scale)

    def standard_f64(self) -> float:
        return lib().fgo_standard_f64(_u64p(self.s))

    def open01_f64(self) -> float:
        return lib().fgo_open01_f64(_u64p(self.s))

    def poisson(self, lam: float) -> float:
        return lib().fgo_poisson_sample(_u64p(self.s), lam)

    def standard_normal(self) -> float:
        return lib().fgo_standard_normal(_u64p(self.s))

    def lognormal(self, mu: float, sigma: float) -> float:
        return lib().fgo_lognormal_sample(_u64p(self.s), mu, sigma)


def zig_tables():
    x = np.zeros(257, dtype=np.float64)
    f = np.zeros(257, dtype=np.float64)
    lib().fgo_zig_tables(x.ctypes.data_as(_F64P), f.ctypes.data_as(_F64P))
    return x, f


def chacha_block(key, counter: int, stream: int, rounds: int) -> np.ndarray:
    k = np.ascontiguousarray(np.array(key, dtype=np.uint32))
    out = np.zeros(16, dtype=np.uint32)
    lib().fgo_chacha_block(k.ctypes.data_as(_U32P), counter, stream, rounds,
                           out.ctypes.data_as(_U32P))
    return out


def stdrng_u64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().fgo_stdrng_u64(seed, n, _u64p(out))
    return out


def make_offsets(seed: int, n: int, sigma: float) -> np.ndarray:
    out = np.zeros((max(n, 1), 2), dtype=np.float32)
    lib().fgo_make_offsets(seed, n, sigma, _f32p(out))
    return out[:n]


def derive(radius_mean, radius_stddev=0.0, dist="const", zoom=1.0, sigma=0.8, n_samples=32,
           max_radius=("quantile", 0.999), cell_delta=None, seed=5489, in_w=1, in_h=1,
           size=None) -> FgoDerived:
    p = FgoParams()
    p.radius_mean = radius_mean
    p.radius_stddev = radius_stddev
    p.dist_kind = 1 if dist == "lognorm" else 0
    p.zoom = zoom
    p.sigma_px = sigma
    p.n_samples = n_samples
    p.max_radius_is_abs = 1 if max_radius[0] == "absolute" else 0
    p.max_radius_value = max_radius[1]
    p.cell_delta = cell_delta if cell_delta is not None else 0.0
    p.seed = seed
    p.in_w = in_w
    p.in_h = in_h
    p.size_w = size[0] if size else 0
    p.size_h = (size[1] or 0) if size else 0
    d = FgoDerived()
    rc = lib().fgo_derive(C.byref(p), C.byref(d))
    if rc != 0:
        raise ValueError("oracle derive failed")
    return d


def plane(d: FgoDerived, n_samples: int, seed: int, zoom: float, dist: str,
          radius_mean: float) -> FgoPlane:
    p = FgoPlane()
    p.in_w, p.in_h, p.out_w, p.out_h = d.in_w, d.in_h, d.out_w, d.out_h
    p.n_samples = n_samples
    p.dist_kind = 1 if dist == "lognorm" else 0
    p.seed = seed
    p.zoom = zoom
    p.delta = d.delta
    p.rm = d.rm
    p.radius_mean = radius_mean
    p.log_mu = d.log_mu
    p.log_sigma = d.log_sigma
    return p


def normalize_plane(u: np.ndarray):
    u = np.ascontiguousarray(u, dtype=np.float32)
    out = np.empty_like(u)
    mx = lib().fgo_normalize_plane(_f32p(u), u.size, _f32p(out))
    return out, mx


def lambda_plane(normalized: np.ndarray, inv_e_pi_r2: float) -> np.ndarray:
    n = np.ascontiguousarray(normalized, dtype=np.float32)
    out = np.empty_like(n)
    lib().fgo_lambda_plane(_f32p(n), n.size, inv_e_pi_r2, _f32p(out))
    return out


def render_pixelwise(p: FgoPlane, lam: np.ndarray, offsets_input: np.ndarray, region=None,
                     threads: int = 0, counters: bool = False):
    lam = np.ascontiguousarray(lam, dtype=np.float32)
    offs = np.ascontiguousarray(offsets_input, dtype=np.float32).reshape(-1, 2)
    assert lam.shape == (p.in_h, p.in_w)
    x0, y0, x1, y1 = region if region else (0, 0, p.out_w, p.out_h)
    out = np.zeros((y1 - y0, x1 - x0), dtype=np.float32)
    cnt = FgoCounters()
    rc = lib().fgo_render_pixelwise(C.byref(p), _f32p(lam), _f32p(offs), x0, y0, x1, y1,
                                    _f32p(out), threads, C.byref(cnt) if counters else None)
    if rc != 0:
        raise ValueError("oracle render_pixelwise failed")
    return (out, cnt) if counters else out


def render_grainwise(p: FgoPlane, lam: np.ndarray, offsets: np.ndarray, threads: int = 0,
                     want_bitsets: bool = False):
    lam = np.ascontiguousarray(lam, dtype=np.float32)
    offs = np.ascontiguousarray(offsets, dtype=np.float32).reshape(-1, 2)
    out = np.zeros((p.out_h, p.out_w), dtype=np.float32)
    lanes = (p.n_samples + 63) // 64
    bits = np.zeros(p.out_w * p.out_h * max(lanes, 1), dtype=np.uint64) if want_bitsets else None
    pairs = C.c_uint64(0)
    rc = lib().fgo_render_grainwise(C.byref(p), _f32p(lam), _f32p(offs), _f32p(out),
                                    _u64p(bits) if bits is not None else None, threads,
                                    C.byref(pairs))
    if rc != 0:
        raise ValueError("oracle render_grainwise failed")
    if want_bitsets:
        return out, bits.reshape(p.out_h, p.out_w, max(lanes, 1)), pairs.value
    return out


def render_grainwise_counted(p: FgoPlane, lam: np.ndarray, offsets: np.ndarray, threads: int = 0):
    """Grain-wise render and its counters (grains, pairs, disk tests, hits)."""
    lam = np.ascontiguousarray(lam, dtype=np.float32)
    offs = np.ascontiguousarray(offsets, dtype=np.float32).reshape(-1, 2)
    out = np.zeros((p.out_h, p.out_w), dtype=np.float32)
    c = FgoGrainCounters()
    rc = lib().fgo_render_grainwise_counted(C.byref(p), _f32p(lam), _f32p(offs), _f32p(out), None, threads,
                                            C.byref(c))
    if rc != 0:
        raise ValueError("oracle render_grainwise failed")
    return out, dict(grains=c.grains, pairs=c.pairs, tests=c.tests, hits=c.hits)


def cell_dump(p: FgoPlane, lam: np.ndarray, i: int, j: int, max_grains: int = 64):
    lam = np.ascontiguousarray(lam, dtype=np.float32)
    h = C.c_uint64(0)
    st = np.zeros(4, dtype=np.uint64)
    g = np.zeros((max_grains, 3), dtype=np.float32)
    q = lib().fgo_cell_dump(C.byref(p), _f32p(lam), i, j, C.byref(h), _u64p(st), _f32p(g),
                            max_grains)
    return h.value, st, q, g[:min(q, max_grains)]


def cell_count(p: FgoPlane, lam: np.ndarray, i: int, j: int) -> int:
    lam = np.ascontiguousarray(lam, dtype=np.float32)
    return int(lib().fgo_cell_count(C.byref(p), _f32p(lam), i, j))


def pixel_dump(p: FgoPlane, lam: np.ndarray, x: int, y: int, max_grains: int = 2048):
    lam = np.ascontiguousarray(lam, dtype=np.float32)
    h = C.c_uint64(0)
    st = np.zeros(4, dtype=np.uint64)
    g = np.zeros((max_grains, 3), dtype=np.float32)
    q = lib().fgo_pixel_dump(C.byref(p), _f32p(lam), x, y, C.byref(h), _u64p(st), _f32p(g),
                             max_grains)
    return h.value, st, q, g[:min(q, max_grains)]


def cell_digests(p: FgoPlane, lam: np.ndarray, origin, size, threads: int = 0):
    """Per cell of the size[0] x size[1] grid from origin: Poisson counts and
    grain digests, as (h, w) arrays."""
    lam = np.ascontiguousarray(lam, dtype=np.float32)
    w, h = int(size[0]), int(size[1])
    q = np.zeros((h, w), dtype=np.uint32)
    d = np.zeros((h, w), dtype=np.uint64)
    rc = lib().fgo_cell_digests(C.byref(p), _f32p(lam), int(origin[0]), int(origin[1]), w, h,
                                q.ctypes.data_as(_U32P), _u64p(d), threads)
    if rc != 0:
        raise ValueError("oracle cell_digests failed")
    return q, d


def pixel_counts(p: FgoPlane, lam: np.ndarray, threads: int = 0, states: bool = False):
    """Every input pixel's grain-wise Poisson count (and the stream state after it)."""
    lam = np.ascontiguousarray(lam, dtype=np.float32)
    q = np.zeros((p.in_h, p.in_w), dtype=np.uint32)
    st = np.zeros((p.in_h, p.in_w, 4), dtype=np.uint64) if states else None
    rc = lib().fgo_pixel_counts(C.byref(p), _f32p(lam), q.ctypes.data_as(_U32P),
                                _u64p(st) if states else None, threads)
    if rc != 0:
        raise ValueError("oracle pixel_counts failed")
    return (q, st) if states else q


def glibc_math(fn: int, x: np.ndarray, threads: int = 0) -> np.ndarray:
    """The C library's f64 math element-wise (fn as FG_MATH_*)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    if lib().fgo_glibc_math(fn, x.ctypes.data_as(_F64P), x.size, out.ctypes.data_as(_F64P), threads) != 0:
        raise ValueError("bad math function")
    return out


def max_threads() -> int:
    return int(lib().fgo_max_threads())


if os.environ.get("FG_ORACLE_REBUILD"):
    build()


# ---- colour workspaces (src/color.rs) ---------------------------------------
COLOR_LUMA, COLOR_RGB = 0, 1


def color_constants() -> np.ndarray:
    """Kr, Kg, Kb and the Cb / Cr denominators as f32 (src/color.rs:9-13)."""
    out = np.zeros(5, dtype=np.float32)
    lib().fgo_color_constants(_f32p(out))
    return out


def to_u8(v) -> np.ndarray:
    """src/color.rs:237-239, element-wise."""
    v = np.asarray(v, dtype=np.float32)
    f = lib().fgo_to_u8
    return np.array([f(float(x)) for x in v.ravel()], dtype=np.uint8).reshape(v.shape)


def color_split(rgb: np.ndarray, mode: int) -> np.ndarray:
    """(h, w, 3) f32 -> (3, h, w): Y, Cb, Cr (mode 0) or R, G, B (mode 1)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    out = np.empty((3, h, w), dtype=np.float32)
    if lib().fgo_color_split(mode, _f32p(rgb), h * w, _f32p(out[0]), _f32p(out[1]), _f32p(out[2])) != 0:
        raise ValueError("oracle color_split failed")
    return out


def resize_nearest(src: np.ndarray, nw: int, nh: int) -> np.ndarray:
    """Plane::resize_nearest (src/model.rs:77-98)."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    h, w = src.shape
    out = np.empty((nh, nw), dtype=np.float32)
    lib().fgo_resize_nearest(_f32p(src), w, h, nw, nh, _f32p(out))
    return out


def color_merge(mode: int, p0: np.ndarray, p1: np.ndarray, p2: np.ndarray) -> np.ndarray:
    """Workspace::into_rgb_image: p0 at (h, w); p1, p2 at (ch, cw) (luma) or
    (h, w) (rgb).  Returns (h, w, 3) u8."""
    p0, p1, p2 = (np.ascontiguousarray(p, dtype=np.float32) for p in (p0, p1, p2))
    h, w = p0.shape
    ch, cw = p1.shape
    out = np.empty((h, w, 3), dtype=np.uint8)
    if lib().fgo_color_merge(mode, _f32p(p0), _f32p(p1), _f32p(p2), w, h, cw, ch,
                             out.ctypes.data_as(C.POINTER(C.c_uint8))) != 0:
        raise ValueError("oracle color_merge failed")
    return out
