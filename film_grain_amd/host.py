"""ctypes binding of the host library's C ABI (include/film_grain.h).

Parameter validation, derived parameters, offsets, the lambda pre-pass and the
library-level render entry points of the C++ host (film_grain_amd/csrc/host),
which mirror the reference's src/params.rs, src/model.rs, src/rng.rs,
src/choose.rs and src/lib.rs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .capi import lib as _lib

FGH_OK = 0
FGH_EPARAMS = 10
FGH_EIMAGE = 11
FGH_EIO = 12
FGH_EGPU = 13
FGH_EUNSUPPORTED = 14
FGH_ECANCELLED = 15
FGH_EMESSAGE = 16
FGH_ECAPACITY = 17

EXPORTED = ["fgh_derive", "fgh_make_offsets", "fgh_normalize_plane", "fgh_lambda_plane",
            "fgh_inverse_normal_cdf", "fgh_render_rgb", "fgh_render_file", "fgh_dry_run",
            "fgh_color_roundtrip", "fgh_load_image", "fgh_save_rgb8"]


class CliArgsC(C.Structure):
    _fields_ = [
        ("input_path", C.c_char_p), ("output_path", C.c_char_p), ("radius_dist", C.c_uint32),
        ("radius_mean", C.c_double), ("radius_stddev", C.c_double), ("zoom", C.c_double),
        ("sigma_px", C.c_double), ("n_samples", C.c_uint32), ("algo", C.c_uint32),
        ("max_radius", C.c_char_p), ("cell", C.c_char_p), ("color_mode", C.c_uint32),
        ("device", C.c_uint32), ("roi", C.c_char_p), ("size", C.c_char_p), ("seed", C.c_uint64),
        ("dry_run", C.c_int32), ("explain", C.c_int32), ("output_format", C.c_char_p),
    ]


class DerivedC(C.Structure):
    _fields_ = [
        ("in_w", C.c_uint32), ("in_h", C.c_uint32), ("out_w", C.c_uint32), ("out_h", C.c_uint32),
        ("n_samples", C.c_uint32), ("algorithm", C.c_uint32), ("dist_kind", C.c_uint32),
        ("seed", C.c_uint64), ("zoom", C.c_float), ("delta", C.c_float), ("rm", C.c_float),
        ("radius_mean", C.c_float), ("radius_stddev", C.c_float), ("log_mu", C.c_float),
        ("log_sigma", C.c_float), ("inv_e_pi_r2", C.c_float), ("sigma_ratio", C.c_float),
        ("rm_ratio", C.c_float),
    ]


class RenderError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


@dataclass
class CliArgs:
    """Mirror of the reference's CliArgs (src/params.rs:93-114) with its CLI
    defaults (src/main.rs:71-247); device defaults to hip here."""
    input_path: str | None = None
    output_path: str | None = None
    radius_dist: str = "const"
    radius: float = 0.10
    radius_stddev: float = 0.0
    zoom: float = 1.0
    sigma: float = 0.8
    iters: int = 32
    algo: str = "auto"
    max_radius: str = "quantile 0.999"
    cell: str = "approx mu_r"
    color_mode: str = "luma"
    device: str = "hip"
    roi: str | None = None
    size: str | None = None
    seed: int = 5489
    dry_run: bool = False
    explain: bool = False
    format: str | None = None

    def to_c(self) -> CliArgsC:
        a = CliArgsC()
        enc = lambda s: s.encode() if s is not None else None
        a.input_path, a.output_path = enc(self.input_path), enc(self.output_path)
        a.radius_dist = {"const": 0, "lognorm": 1}[self.radius_dist]
        a.radius_mean, a.radius_stddev = self.radius, self.radius_stddev
        a.zoom, a.sigma_px = self.zoom, self.sigma
        a.n_samples = self.iters
        a.algo = {"auto": 0, "grain": 1, "pixel": 2}[self.algo]
        a.max_radius, a.cell = enc(self.max_radius), enc(self.cell)
        a.color_mode = {"luma": 0, "rgb": 1}[self.color_mode]
        a.device = {"cpu": 0, "gpu": 1, "hip": 2}[self.device]
        a.roi, a.size = enc(self.roi), enc(self.size)
        a.seed = self.seed
        a.dry_run, a.explain = int(self.dry_run), int(self.explain)
        a.output_format = enc(self.format)
        return a


def _setup():
    L = _lib()
    if getattr(L, "_fgh_ready", False):
        return L
    fp = C.POINTER(C.c_float)
    L.fgh_derive.argtypes = [C.POINTER(CliArgsC), C.c_uint32, C.c_uint32, C.POINTER(DerivedC), fp, fp,
                             C.c_char_p, C.c_size_t]
    L.fgh_derive.restype = C.c_int
    L.fgh_make_offsets.argtypes = [C.c_uint64, C.c_uint32, C.c_float, fp]
    L.fgh_make_offsets.restype = C.c_int
    L.fgh_normalize_plane.argtypes = [fp, C.c_size_t, fp]
    L.fgh_normalize_plane.restype = C.c_float
    L.fgh_lambda_plane.argtypes = [fp, C.c_size_t, C.c_float, fp]
    L.fgh_lambda_plane.restype = None
    L.fgh_inverse_normal_cdf.argtypes = [C.c_double]
    L.fgh_inverse_normal_cdf.restype = C.c_double
    L.fgh_render_rgb.argtypes = [C.POINTER(CliArgsC), fp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8),
                                 C.c_size_t, fp, C.c_size_t, C.POINTER(DerivedC), C.CFUNCTYPE(C.c_int, C.c_void_p),
                                 C.c_void_p, C.c_char_p, C.c_size_t]
    L.fgh_render_rgb.restype = C.c_int
    L.fgh_render_rgb_frame.argtypes = L.fgh_render_rgb.argtypes
    L.fgh_render_rgb_frame.restype = C.c_int
    L.fgh_logf_array.argtypes = [fp, C.c_size_t, fp]
    L.fgh_logf_array.restype = None
    L.fgh_color_split.argtypes = [fp, C.c_uint32, C.c_uint32, C.c_uint32, fp, fp, fp]
    L.fgh_color_split.restype = C.c_int
    L.fgh_render_file.argtypes = [C.POINTER(CliArgsC), C.POINTER(DerivedC), C.c_char_p, C.c_size_t]
    L.fgh_render_file.restype = C.c_int
    L.fgh_dry_run.argtypes = [C.POINTER(CliArgsC), C.POINTER(DerivedC), C.c_char_p, C.c_size_t]
    L.fgh_dry_run.restype = C.c_int
    L.fgh_color_roundtrip.argtypes = [fp, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8)]
    L.fgh_color_roundtrip.restype = C.c_int
    L.fgh_load_image.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), fp, C.c_size_t,
                                 C.c_char_p, C.c_size_t]
    L.fgh_load_image.restype = C.c_int
    L.fgh_save_rgb8.argtypes = [C.c_char_p, C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32, C.c_char_p,
                                C.c_char_p, C.c_size_t]
    L.fgh_save_rgb8.restype = C.c_int
    L._fgh_ready = True
    return L


def color_roundtrip(rgb: np.ndarray, color_mode: str) -> np.ndarray:
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    out = np.zeros((h, w, 3), dtype=np.uint8)
    rc = _setup().fgh_color_roundtrip(_f32(rgb), w, h, 1 if color_mode == "rgb" else 0,
                                      out.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc != FGH_OK:
        raise RenderError(rc, "colour round trip failed")
    return out


def load_image(path: str) -> np.ndarray:
    L = _setup()
    w, h = C.c_uint32(0), C.c_uint32(0)
    err = C.create_string_buffer(512)
    rc = L.fgh_load_image(path.encode(), C.byref(w), C.byref(h), None, 0, err, 512)
    if rc != FGH_OK:
        raise RenderError(rc, err.value.decode())
    out = np.zeros((h.value, w.value, 3), dtype=np.float32)
    rc = L.fgh_load_image(path.encode(), C.byref(w), C.byref(h), _f32(out), out.size, err, 512)
    if rc != FGH_OK:
        raise RenderError(rc, err.value.decode())
    return out


def save_rgb8(path: str, rgb8: np.ndarray, fmt: str = "png") -> None:
    rgb8 = np.ascontiguousarray(rgb8, dtype=np.uint8)
    h, w = rgb8.shape[:2]
    err = C.create_string_buffer(512)
    rc = _setup().fgh_save_rgb8(path.encode(), rgb8.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, fmt.encode(),
                                err, 512)
    if rc != FGH_OK:
        raise RenderError(rc, err.value.decode())


def _f32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def derive(args: CliArgs, in_w: int, in_h: int):
    """Returns (DerivedC, offsets (N,2), offsets_input (N,2)); raises RenderError."""
    L = _setup()
    d = DerivedC()
    n = max(int(args.iters), 1)
    offs = np.zeros((n, 2), dtype=np.float32)
    offs_in = np.zeros((n, 2), dtype=np.float32)
    err = C.create_string_buffer(512)
    rc = L.fgh_derive(C.byref(args.to_c()), in_w, in_h, C.byref(d), _f32(offs), _f32(offs_in), err, 512)
    if rc != FGH_OK:
        raise RenderError(rc, err.value.decode())
    return d, offs, offs_in


def make_offsets(seed: int, n: int, sigma: float) -> np.ndarray:
    out = np.zeros((max(n, 1), 2), dtype=np.float32)
    _setup().fgh_make_offsets(seed, n, sigma, _f32(out))
    return out[:n]


def normalize_plane(u: np.ndarray):
    u = np.ascontiguousarray(u, dtype=np.float32)
    out = np.empty_like(u)
    mx = _setup().fgh_normalize_plane(_f32(u), u.size, _f32(out))
    return out, mx


def lambda_plane(normalized: np.ndarray, inv_e_pi_r2: float) -> np.ndarray:
    n = np.ascontiguousarray(normalized, dtype=np.float32)
    out = np.empty_like(n)
    _setup().fgh_lambda_plane(_f32(n), n.size, inv_e_pi_r2, _f32(out))
    return out


def logf_array(x: np.ndarray) -> np.ndarray:
    """The C library's logf, elementwise (the reference's f32::ln)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    _setup().fgh_logf_array(_f32(x), x.size, _f32(out))
    return out


def color_split(rgb: np.ndarray, color_mode: str) -> np.ndarray:
    """Workspace planes of an (h, w, 3) f32 image: (3, h, w) as Y, Cb, Cr
    (luma) or R, G, B (rgb)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    out = np.empty((3, h, w), dtype=np.float32)
    rc = _setup().fgh_color_split(_f32(rgb), w, h, 1 if color_mode == "rgb" else 0, _f32(out[0]), _f32(out[1]),
                                  _f32(out[2]))
    if rc != FGH_OK:
        raise RenderError(rc, "colour split failed")
    return out


def inverse_normal_cdf(p: float) -> float:
    return _setup().fgh_inverse_normal_cdf(p)


def render_rgb(args: CliArgs, rgb: np.ndarray, cancel=None, frame_on_device: bool = False,
               with_planes: bool = True):
    """Library-level render of an (h, w, 3) f32 image; returns (rgb8, planes,
    stats) (planes None without with_planes).  frame_on_device: the colour
    split, lambda pre-pass and merge run on the device too (one upload, one
    download); the result is the same."""
    L = _setup()
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    d, _, _ = derive(args, w, h) if args.roi is None else (None, None, None)
    if d is None:  # ROI: size after the crop
        x0, y0 = (int(v) for v in args.roi.split(":")[0].split(","))
        x1, y1 = (int(v) for v in args.roi.split(":")[1].split(","))
        d, _, _ = derive(args, x1 - x0, y1 - y0)
    nplanes = 3 if args.color_mode == "rgb" else 1
    out8 = np.zeros((d.out_h, d.out_w, 3), dtype=np.uint8)
    planes = np.zeros((nplanes, d.out_h, d.out_w), dtype=np.float32) if with_planes else None
    stats = DerivedC()
    err = C.create_string_buffer(512)
    cb_type = C.CFUNCTYPE(C.c_int, C.c_void_p)
    cb = cb_type(lambda _u: 1 if cancel() else 0) if cancel else cb_type(0)
    fn = L.fgh_render_rgb_frame if frame_on_device else L.fgh_render_rgb
    rc = fn(C.byref(args.to_c()), _f32(rgb), w, h, out8.ctypes.data_as(C.POINTER(C.c_uint8)),
            out8.size, _f32(planes) if with_planes else None, planes.size if with_planes else 0, C.byref(stats),
            cb, None, err, 512)
    if rc != FGH_OK:
        raise RenderError(rc, err.value.decode())
    return out8, planes, stats


def render_file(args: CliArgs):
    L = _setup()
    stats = DerivedC()
    err = C.create_string_buffer(512)
    rc = L.fgh_render_file(C.byref(args.to_c()), C.byref(stats), err, 512)
    if rc != FGH_OK:
        raise RenderError(rc, err.value.decode())
    return stats


def dry_run(args: CliArgs):
    L = _setup()
    stats = DerivedC()
    err = C.create_string_buffer(512)
    rc = L.fgh_dry_run(C.byref(args.to_c()), C.byref(stats), err, 512)
    if rc != FGH_OK:
        raise RenderError(rc, err.value.decode())
    return stats
