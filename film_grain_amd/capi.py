"""ctypes binding of libfilmgrain_hip.so (include/film_grain_hip.h).

Loads the in-tree library built by film_grain_amd.build and raises if it is
missing: there is no fallback path.  Device buffers are torch tensors (HIP
memory) passed by data_ptr(); streams are torch.cuda streams.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libfilmgrain_hip.so"

FG_OK = 0
FG_EINVAL = 1
FG_ENODEV = 2
FG_ENOMEM = 3
FG_ECANCELLED = 4
FG_EHIP = 5
FG_ENEIGHBOURHOOD = 6

ERROR_NAMES = {
    FG_EINVAL: "FG_EINVAL", FG_ENODEV: "FG_ENODEV", FG_ENOMEM: "FG_ENOMEM",
    FG_ECANCELLED: "FG_ECANCELLED", FG_EHIP: "FG_EHIP", FG_ENEIGHBOURHOOD: "FG_ENEIGHBOURHOOD",
}

# Every symbol include/film_grain_hip.h declares.
EXPORTED = [
    "fg_ctx_create", "fg_ctx_destroy", "fg_last_error", "fg_render_pixelwise",
    "fg_render_grainwise", "fg_render_pixelwise_device", "fg_render_grainwise_device",
    "fg_stream_status", "fg_pixelwise_stats_device", "fg_grainwise_stats_device",
    "fg_set_launch_config", "fg_get_launch_config", "fg_set_timing", "fg_get_timings",
    "fg_batch_create", "fg_batch_run", "fg_batch_status", "fg_batch_info", "fg_batch_destroy",
    "fg_mlp_features_device", "fg_mlp_forward_device", "fg_mlp_grad_device", "fg_mlp_adam_device",
    "fg_mlp_budget_device", "fg_mlp_budget_scale_device", "fg_budget_scale_device",
    "fg_render_pixelwise_adaptive_device", "fg_render_frame_device", "fg_render_frame",
    "fg_lambda_device", "fg_color_split_device", "fg_color_merge_device", "fg_log_f32_device", "fg_cancel",
    "fg_pixelwise_grains", "fg_grainwise_pixel_state", "fg_max_grains", "fg_render_pixelwise_device_noalloc",
    "fg_math_device", "fg_launch_bucket", "fg_set_launch_table", "fg_get_launch_table",
    "fg_render_pixelwise_planes_device",
    "fg_version",
]
COLOR_LUMA, COLOR_RGB = 0, 1
ALGO_PIXEL, ALGO_GRAIN = 0, 1
MLP_PARAMS = 5313
MLP_FEATURES = 16
LOSS_SQUARED = 0
LOSS_NLL = 1
FG_ECAPACITY = 7
ERROR_NAMES[FG_ECAPACITY] = "FG_ECAPACITY"
NOALLOC_SCRATCH_BYTES = 256 << 20
MATH_EXP_NEG, MATH_EXP, MATH_LOG, MATH_TAN_PI, MATH_LOG_GAMMA, MATH_SQRT, MATH_EXP_F32 = range(7)


class PlaneDesc(C.Structure):
    _fields_ = [
        ("in_w", C.c_uint32), ("in_h", C.c_uint32), ("out_w", C.c_uint32), ("out_h", C.c_uint32),
        ("n_samples", C.c_uint32), ("dist_kind", C.c_uint32), ("seed", C.c_uint64),
        ("zoom", C.c_float), ("delta", C.c_float), ("rm", C.c_float),
        ("radius_mean", C.c_float), ("log_mu", C.c_float), ("log_sigma", C.c_float),
    ]


class FrameDesc(C.Structure):
    _fields_ = [("plane", PlaneDesc), ("color_mode", C.c_uint32), ("algo", C.c_uint32),
                ("inv_e_pi_r2", C.c_float), ("reserved", C.c_uint32)]


class LaunchConfig(C.Structure):
    _fields_ = [
        ("pix_tile_w", C.c_int), ("pix_tile_h", C.c_int), ("pix_lds_bytes", C.c_int),
        ("grain_tile_w", C.c_int), ("grain_tile_h", C.c_int), ("grain_lds_bytes", C.c_int),
        ("grain_grains_per_thread", C.c_int), ("pix_variant", C.c_int), ("pix_block", C.c_int),
        ("grain_variant", C.c_int), ("grain_block", C.c_int),
    ]


class LaunchKey(C.Structure):
    _fields_ = [("algo", C.c_uint32), ("dist_kind", C.c_uint32), ("zoom_q", C.c_int32), ("ratio_h", C.c_int32),
                ("n_bucket", C.c_uint32)]


class LaunchEntry(C.Structure):
    _fields_ = [("key", LaunchKey), ("cfg", LaunchConfig)]


CANCEL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)


class FilmGrainError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Loads the in-tree HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m film_grain_amd.build` "
            "(the HIP backend has no CPU fallback)")
    # PyTorch ships its own HIP runtime under the same soname
    # (libamdhip64.so.7).  Loaded first, it satisfies this library's
    # dependency, so both share one runtime and device pointers pass freely;
    # loaded second, it would bring a second runtime into the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    fp = C.POINTER(C.c_float)
    L.fg_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.fg_ctx_create.restype = C.c_int
    L.fg_ctx_destroy.argtypes = [vp]
    L.fg_ctx_destroy.restype = None
    L.fg_last_error.argtypes = [vp]
    L.fg_last_error.restype = C.c_char_p
    for name in ("fg_render_pixelwise", "fg_render_grainwise"):
        f = getattr(L, name)
        f.argtypes = [vp, C.POINTER(PlaneDesc), fp, fp, fp, CANCEL_FN, vp]
        f.restype = C.c_int
    for name in ("fg_render_pixelwise_device", "fg_render_grainwise_device"):
        f = getattr(L, name)
        f.argtypes = [vp, C.POINTER(PlaneDesc), vp, vp, vp, vp]
        f.restype = C.c_int
    L.fg_stream_status.argtypes = [vp, vp]
    L.fg_stream_status.restype = C.c_int
    for name in ("fg_pixelwise_stats_device", "fg_grainwise_stats_device"):
        f = getattr(L, name)
        f.argtypes = [vp, C.POINTER(PlaneDesc), vp, vp, vp, C.POINTER(C.c_uint64), vp]
        f.restype = C.c_int
    L.fg_set_launch_config.argtypes = [vp, C.POINTER(LaunchConfig)]
    L.fg_set_launch_config.restype = C.c_int
    L.fg_get_launch_config.argtypes = [vp, C.POINTER(PlaneDesc), C.c_int, C.POINTER(LaunchConfig)]
    L.fg_get_launch_config.restype = C.c_int
    L.fg_render_pixelwise_planes_device.argtypes = [vp, C.POINTER(PlaneDesc), C.c_int, C.POINTER(vp), vp,
                                                    C.POINTER(vp), vp]
    L.fg_render_pixelwise_planes_device.restype = C.c_int
    L.fg_launch_bucket.argtypes = [C.POINTER(PlaneDesc), C.c_int, C.POINTER(LaunchKey)]
    L.fg_launch_bucket.restype = C.c_int
    L.fg_set_launch_table.argtypes = [vp, C.POINTER(LaunchEntry), C.c_int]
    L.fg_set_launch_table.restype = C.c_int
    L.fg_get_launch_table.argtypes = [vp, C.POINTER(LaunchEntry), C.c_int]
    L.fg_get_launch_table.restype = C.c_int
    L.fg_set_timing.argtypes = [vp, C.c_int]
    L.fg_set_timing.restype = C.c_int
    L.fg_get_timings.argtypes = [vp, C.POINTER(C.c_float)]
    L.fg_get_timings.restype = C.c_int
    L.fg_batch_create.argtypes = [vp, C.POINTER(PlaneDesc), C.c_int, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                  C.c_int, vp, C.POINTER(vp)]
    L.fg_batch_create.restype = C.c_int
    for name in ("fg_batch_run", "fg_batch_status"):
        f = getattr(L, name)
        f.argtypes = [vp, vp]
        f.restype = C.c_int
    L.fg_batch_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fg_batch_info.restype = C.c_int
    L.fg_batch_destroy.argtypes = [vp]
    L.fg_batch_destroy.restype = None
    u32 = C.c_uint32
    L.fg_mlp_features_device.argtypes = [vp, vp, u32, u32, u32, u32, C.c_float, C.c_float, C.c_float, C.c_float,
                                         C.c_int, vp, vp]
    L.fg_mlp_forward_device.argtypes = [vp, vp, vp, C.c_int64, vp, vp]
    L.fg_mlp_grad_device.argtypes = [vp, vp, vp, vp, C.c_int64, C.c_int64, C.c_int, vp, vp, vp]
    L.fg_mlp_adam_device.argtypes = [vp, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, vp]
    L.fg_mlp_budget_device.argtypes = [vp, vp, vp, C.c_int64, C.c_float, C.c_int, C.c_int, vp, vp]
    L.fg_mlp_budget_scale_device.argtypes = [vp, vp, vp, C.c_int64, vp, C.c_int, C.c_int, vp, vp]
    L.fg_budget_scale_device.argtypes = [vp, vp, C.c_int64, C.c_float, C.c_double, C.c_int, C.c_int, C.c_int, vp,
                                         vp]
    L.fg_render_pixelwise_adaptive_device.argtypes = [vp, C.POINTER(PlaneDesc), vp, vp, vp, vp, vp]
    L.fg_render_frame_device.argtypes = [vp, C.POINTER(FrameDesc), vp, vp, vp, vp, vp]
    L.fg_render_frame.argtypes = [vp, C.POINTER(FrameDesc), fp, fp, C.POINTER(C.c_uint8), fp, CANCEL_FN, vp]
    L.fg_lambda_device.argtypes = [vp, vp, u32, u32, C.c_float, vp, vp, vp]
    L.fg_color_split_device.argtypes = [vp, C.c_int, vp, u32, u32, vp, vp, vp, vp]
    L.fg_color_merge_device.argtypes = [vp, C.c_int, vp, vp, vp, u32, u32, u32, u32, vp, vp]
    L.fg_log_f32_device.argtypes = [vp, vp, C.c_int64, vp, vp]
    L.fg_cancel.argtypes = [vp, C.c_int]
    L.fg_cancel.restype = C.c_int
    for name in ("fg_mlp_features_device", "fg_mlp_forward_device", "fg_mlp_grad_device", "fg_mlp_adam_device",
                 "fg_mlp_budget_device", "fg_mlp_budget_scale_device", "fg_budget_scale_device",
                 "fg_render_pixelwise_adaptive_device", "fg_render_frame_device",
                 "fg_render_frame", "fg_lambda_device", "fg_color_split_device", "fg_color_merge_device",
                 "fg_log_f32_device"):
        getattr(L, name).restype = C.c_int
    L.fg_version.argtypes = []
    L.fg_version.restype = C.c_char_p
    L.fg_pixelwise_grains.argtypes = [vp, vp, vp, vp, vp, C.c_uint64, vp, vp, vp]
    L.fg_pixelwise_grains.restype = C.c_int
    L.fg_grainwise_pixel_state.argtypes = [vp, vp, vp, vp, vp]
    L.fg_grainwise_pixel_state.restype = C.c_int
    L.fg_render_pixelwise_device_noalloc.argtypes = [vp, vp, vp, vp, vp, vp]
    L.fg_render_pixelwise_device_noalloc.restype = C.c_int
    L.fg_math_device.argtypes = [vp, C.c_int, vp, C.c_int64, vp, vp]
    L.fg_math_device.restype = C.c_int
    L.fg_max_grains.argtypes = []
    L.fg_max_grains.restype = C.c_uint64
    _lib = L
    return L


def make_desc(in_w, in_h, out_w, out_h, n_samples, dist_kind, seed, zoom, delta, rm, radius_mean,
              log_mu=0.0, log_sigma=0.0) -> PlaneDesc:
    d = PlaneDesc()
    d.in_w, d.in_h, d.out_w, d.out_h = in_w, in_h, out_w, out_h
    d.n_samples, d.dist_kind, d.seed = n_samples, dist_kind, seed
    d.zoom, d.delta, d.rm, d.radius_mean = zoom, delta, rm, radius_mean
    d.log_mu, d.log_sigma = log_mu, log_sigma
    return d


class Context:
    """Owns one fg_ctx (one HIP device)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.fg_ctx_create(device, C.byref(h))
        if rc != FG_OK:
            raise FilmGrainError(rc, "no compatible GPU adapter found" if rc == FG_ENODEV
                                 else "context creation failed")
        self.handle = h

    def close(self):
        if self.handle:
            self._L.fg_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self) -> str:
        return self._L.fg_last_error(self.handle).decode()

    def _check(self, rc: int):
        if rc != FG_OK:
            raise FilmGrainError(rc, self.last_error())

    # ---- host-memory entry points (numpy f32 arrays) ----
    def render_host(self, algo: str, desc: PlaneDesc, lam, offsets, cancel=None):
        import numpy as np
        lam = np.ascontiguousarray(lam, dtype=np.float32)
        offs = np.ascontiguousarray(offsets, dtype=np.float32)
        out = np.empty((desc.out_h, desc.out_w), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        cb = CANCEL_FN(lambda _u: 1 if cancel() else 0) if cancel else CANCEL_FN(0)
        fn = self._L.fg_render_pixelwise if algo == "pixel" else self._L.fg_render_grainwise
        rc = fn(self.handle, C.byref(desc), lam.ctypes.data_as(fp), offs.ctypes.data_as(fp),
                out.ctypes.data_as(fp), cb, None)
        self._check(rc)
        return out

    # ---- device-memory entry points (torch tensors on the HIP device) ----
    def render_device(self, algo: str, desc: PlaneDesc, lam, offsets, out, stream=None):
        fn = self._L.fg_render_pixelwise_device if algo == "pixel" else self._L.fg_render_grainwise_device
        rc = fn(self.handle, C.byref(desc), C.c_void_p(lam.data_ptr()), C.c_void_p(offsets.data_ptr()),
                C.c_void_p(out.data_ptr()), C.c_void_p(_stream_handle(stream)))
        self._check(rc)

    def render_planes_device(self, desc: PlaneDesc, lams, offsets, outs, stream=None):
        """The pixel-wise planes of one frame (1-3 lambda tensors, one
        description): the cells' random streams are hashed once for all of
        them (fg_render_pixelwise_planes_device)."""
        n = len(lams)
        if n != len(outs):
            raise ValueError("one output per lambda plane")
        la = (C.c_void_p * n)(*[l.data_ptr() for l in lams])
        oa = (C.c_void_p * n)(*[o.data_ptr() for o in outs])
        rc = self._L.fg_render_pixelwise_planes_device(self.handle, C.byref(desc), n, la,
                                                       C.c_void_p(offsets.data_ptr()), oa,
                                                       C.c_void_p(_stream_handle(stream)))
        self._check(rc)

    def render_device_noalloc(self, desc: PlaneDesc, lam, offsets, out, stream=None):
        """Pixel-wise render like render_device, with every scratch buffer in
        the library's static device array (fg_render_pixelwise_device_noalloc)."""
        rc = self._L.fg_render_pixelwise_device_noalloc(
            self.handle, C.byref(desc), C.c_void_p(lam.data_ptr()), C.c_void_p(offsets.data_ptr()),
            C.c_void_p(out.data_ptr()), C.c_void_p(_stream_handle(stream)))
        self._check(rc)

    def render_adaptive_device(self, desc: PlaneDesc, lam, offsets, budget, out, stream=None):
        """Pixel-wise render where output pixel i uses the first budget[i]
        (u16) of the desc.n_samples offsets."""
        rc = self._L.fg_render_pixelwise_adaptive_device(
            self.handle, C.byref(desc), C.c_void_p(lam.data_ptr()), C.c_void_p(offsets.data_ptr()),
            C.c_void_p(budget.data_ptr()), C.c_void_p(out.data_ptr()), C.c_void_p(_stream_handle(stream)))
        self._check(rc)

    # ---- frames: colour split, lambda pre-pass, render, merge ----
    def render_frame_device(self, fd: FrameDesc, rgb, offsets, rgb8, planes=None, stream=None):
        """rgb: (in_h, in_w, 3) f32 tensor; rgb8: (out_h, out_w, 3) uint8
        tensor; planes: optional (k, out_h, out_w) f32 tensor."""
        self._check(self._L.fg_render_frame_device(
            self.handle, C.byref(fd), C.c_void_p(rgb.data_ptr()), C.c_void_p(offsets.data_ptr()),
            C.c_void_p(rgb8.data_ptr()), C.c_void_p(planes.data_ptr() if planes is not None else 0),
            C.c_void_p(_stream_handle(stream))))

    def render_frame(self, fd: FrameDesc, rgb: np.ndarray, offsets: np.ndarray, with_planes: bool = False,
                     cancel=None):
        """Host-memory frame: returns (rgb8 (out_h, out_w, 3), planes or None)."""
        d = fd.plane
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        offs = np.ascontiguousarray(offsets, dtype=np.float32)
        out8 = np.empty((d.out_h, d.out_w, 3), dtype=np.uint8)
        nr = 3 if fd.color_mode == COLOR_RGB else 1
        planes = np.empty((nr, d.out_h, d.out_w), dtype=np.float32) if with_planes else None
        fp = C.POINTER(C.c_float)
        cb = CANCEL_FN(lambda _u: 1 if cancel() else 0) if cancel else CANCEL_FN(0)
        self._check(self._L.fg_render_frame(
            self.handle, C.byref(fd), rgb.ctypes.data_as(fp), offs.ctypes.data_as(fp),
            out8.ctypes.data_as(C.POINTER(C.c_uint8)), planes.ctypes.data_as(fp) if planes is not None else None,
            cb, None))
        return out8, planes

    def lambda_device(self, plane, inv_e_pi_r2, lam, norm=None, stream=None):
        h, w = plane.shape
        self._check(self._L.fg_lambda_device(
            self.handle, C.c_void_p(plane.data_ptr()), w, h, inv_e_pi_r2,
            C.c_void_p(norm.data_ptr() if norm is not None else 0), C.c_void_p(lam.data_ptr()),
            C.c_void_p(_stream_handle(stream))))

    def color_split_device(self, mode: int, rgb, p0, p1, p2, stream=None):
        h, w = rgb.shape[:2]
        self._check(self._L.fg_color_split_device(
            self.handle, mode, C.c_void_p(rgb.data_ptr()), w, h, C.c_void_p(p0.data_ptr()),
            C.c_void_p(p1.data_ptr()), C.c_void_p(p2.data_ptr()), C.c_void_p(_stream_handle(stream))))

    def color_merge_device(self, mode: int, p0, p1, p2, rgb8, stream=None):
        h, w = p0.shape
        ch, cw = p1.shape
        self._check(self._L.fg_color_merge_device(
            self.handle, mode, C.c_void_p(p0.data_ptr()), C.c_void_p(p1.data_ptr()), C.c_void_p(p2.data_ptr()),
            w, h, cw, ch, C.c_void_p(rgb8.data_ptr()), C.c_void_p(_stream_handle(stream))))

    def log_f32_device(self, x, out, stream=None):
        self._check(self._L.fg_log_f32_device(self.handle, C.c_void_p(x.data_ptr()), x.numel(),
                                              C.c_void_p(out.data_ptr()), C.c_void_p(_stream_handle(stream))))

    # ---- generator inspection (parity tests) ----
    def pixelwise_grains(self, desc: PlaneDesc, lam, offsets):
        """Every grain record of the plane's HBM generation, as (n, 4) f32
        (cx, cy, r, cell bits), with the extended grid's origin and size."""
        import numpy as np
        n = C.c_uint64(0)
        origin = (C.c_int32 * 2)()
        size = (C.c_uint32 * 2)()
        args = (C.byref(desc), C.c_void_p(lam.data_ptr()), C.c_void_p(offsets.data_ptr()))
        rc = self._L.fg_pixelwise_grains(self.handle, *args, None, 0, C.byref(n), origin, size)
        if rc not in (FG_OK, FG_ECAPACITY):
            self._check(rc)
        out = np.zeros((max(int(n.value), 1), 4), dtype=np.float32)
        rc = self._L.fg_pixelwise_grains(self.handle, *args, C.c_void_p(out.ctypes.data), out.shape[0],
                                         C.byref(n), origin, size)
        self._check(rc)
        return out[:int(n.value)], (origin[0], origin[1]), (size[0], size[1])

    def grainwise_pixel_state(self, desc: PlaneDesc, lam):
        """Per input pixel: the grain count and the Xoshiro state after the count draws."""
        import numpy as np
        npx = desc.in_w * desc.in_h
        q = np.zeros(npx, dtype=np.uint32)
        st = np.zeros((npx, 4), dtype=np.uint64)
        self._check(self._L.fg_grainwise_pixel_state(self.handle, C.byref(desc), C.c_void_p(lam.data_ptr()),
                                                     C.c_void_p(q.ctypes.data), C.c_void_p(st.ctypes.data)))
        return q.reshape(desc.in_h, desc.in_w), st.reshape(desc.in_h, desc.in_w, 4)

    def math_device(self, fn: int, x, out, stream=None):
        """The integrators' f64 math on the device (FG_MATH_* codes), element-wise."""
        self._check(self._L.fg_math_device(self.handle, fn, C.c_void_p(x.data_ptr()), x.numel(),
                                           C.c_void_p(out.data_ptr()), C.c_void_p(_stream_handle(stream))))

    def cancel(self, on: bool = True):
        """Raise (or lower) the context's cancel flag: the integrators' tiles
        not yet started return at once and the status reports FG_ECANCELLED."""
        self._check(self._L.fg_cancel(self.handle, 1 if on else 0))

    def stream_status(self, stream=None):
        self._check(self._L.fg_stream_status(self.handle, C.c_void_p(_stream_handle(stream))))

    def stats_device(self, algo: str, desc: PlaneDesc, lam, offsets, out, stream=None):
        """Instrumented render.  Pixel-wise: [points, disk tests T, hits];
        grain-wise: [grains, (grain, sample) pairs, disk tests, hits]."""
        fn = self._L.fg_pixelwise_stats_device if algo == "pixel" else self._L.fg_grainwise_stats_device
        cnt = (C.c_uint64 * (3 if algo == "pixel" else 4))()
        rc = fn(self.handle, C.byref(desc), C.c_void_p(lam.data_ptr()), C.c_void_p(offsets.data_ptr()),
                C.c_void_p(out.data_ptr()), cnt, C.c_void_p(_stream_handle(stream)))
        self._check(rc)
        return [int(v) for v in cnt]

    # ---- per-pixel error model (torch tensors on the device) ----
    def mlp_features(self, u, out_w, out_h, zoom, radius, radius_sd, sigma, channel, out=None, stream=None):
        """BF16 features (as int16 tensor [out_h*out_w, 16]) of a normalised
        intensity plane u (f32 [in_h, in_w])."""
        import torch
        in_h, in_w = u.shape
        if out is None:
            out = torch.empty((out_h * out_w, MLP_FEATURES), dtype=torch.int16, device=u.device)
        self._check(self._L.fg_mlp_features_device(
            self.handle, C.c_void_p(u.data_ptr()), in_w, in_h, out_w, out_h, zoom, radius, radius_sd, sigma,
            channel, C.c_void_p(out.data_ptr()), C.c_void_p(_stream_handle(stream))))
        return out

    def mlp_forward(self, params, feats, out=None, stream=None):
        import torch
        n = feats.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=feats.device)
        self._check(self._L.fg_mlp_forward_device(self.handle, C.c_void_p(params.data_ptr()),
                                                  C.c_void_p(feats.data_ptr()), n, C.c_void_p(out.data_ptr()),
                                                  C.c_void_p(_stream_handle(stream))))
        return out

    def mlp_grad(self, params, feats, target, grad, loss, n_total=None, loss_kind="squared", stream=None):
        """Adds the gradient of the mean loss ("squared": (y - target)^2,
        "nll": Gaussian negative log-likelihood of the residual target under
        the predicted log std y) to grad and the loss sum to loss."""
        n = feats.shape[0]
        kind = {"squared": LOSS_SQUARED, "nll": LOSS_NLL}[loss_kind]
        self._check(self._L.fg_mlp_grad_device(self.handle, C.c_void_p(params.data_ptr()),
                                               C.c_void_p(feats.data_ptr()), C.c_void_p(target.data_ptr()), n,
                                               n_total or n, kind, C.c_void_p(grad.data_ptr()),
                                               C.c_void_p(loss.data_ptr()), C.c_void_p(_stream_handle(stream))))

    def mlp_adam(self, params, grad, m, v, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        self._check(self._L.fg_mlp_adam_device(self.handle, C.c_void_p(params.data_ptr()), C.c_void_p(grad.data_ptr()),
                                               C.c_void_p(m.data_ptr()), C.c_void_p(v.data_ptr()), lr, beta1, beta2,
                                               eps, step, C.c_void_p(_stream_handle(stream))))

    def mlp_budget(self, pred, feats, n, c, n_min, n_max, out=None, stream=None):
        """u16 sample budgets (int16 tensor) clamp(ceil(c * s), n_min, n_max):
        s = exp(pred) or, with pred None, sqrt(u (1 - u)) from the features."""
        import torch
        dev = (pred if pred is not None else feats).device
        if out is None:
            out = torch.empty(n, dtype=torch.int16, device=dev)
        self._check(self._L.fg_mlp_budget_device(
            self.handle, C.c_void_p(pred.data_ptr() if pred is not None else 0),
            C.c_void_p(feats.data_ptr() if feats is not None else 0), n, c, n_min, n_max,
            C.c_void_p(out.data_ptr()), C.c_void_p(_stream_handle(stream))))
        return out

    def budget_scale(self, pred, s_mult, target_mse, n_min, n_max, n_ref, out=None, stream=None):
        """The frame's budget scale searched on the device (no host round
        trip): out[0] = c for s = exp(pred) * s_mult, out[1] = c * s_mult.
        pred is one float32 tensor over all the frame's rows."""
        import torch
        if out is None:
            out = torch.empty(2, dtype=torch.float32, device=pred.device)
        self._check(self._L.fg_budget_scale_device(
            self.handle, C.c_void_p(pred.data_ptr()), pred.numel(), s_mult, target_mse, n_min, n_max, n_ref,
            C.c_void_p(out.data_ptr()), C.c_void_p(_stream_handle(stream))))
        return out

    def mlp_budget_scale(self, pred, feats, n, d_c, n_min, n_max, out=None, stream=None):
        """mlp_budget with the scale read from device memory (d_c: a
        one-element float32 tensor, e.g. budget_scale(...)[1:])."""
        import torch
        dev = (pred if pred is not None else feats).device
        if out is None:
            out = torch.empty(n, dtype=torch.int16, device=dev)
        self._check(self._L.fg_mlp_budget_scale_device(
            self.handle, C.c_void_p(pred.data_ptr() if pred is not None else 0),
            C.c_void_p(feats.data_ptr() if feats is not None else 0), n, C.c_void_p(d_c.data_ptr()), n_min, n_max,
            C.c_void_p(out.data_ptr()), C.c_void_p(_stream_handle(stream))))
        return out

    def set_timing(self, enable: bool):
        self._check(self._L.fg_set_timing(self.handle, 1 if enable else 0))

    def timings(self):
        """(whole render, generation kernels, coverage-test kernel) in ms for
        the last render since timing was enabled."""
        ms = (C.c_float * 3)()
        self._check(self._L.fg_get_timings(self.handle, ms))
        return tuple(float(v) for v in ms)

    def set_launch_config(self, **kw):
        cfg = LaunchConfig()
        for k, v in kw.items():
            setattr(cfg, k, int(v))
        self._check(self._L.fg_set_launch_config(self.handle, C.byref(cfg)))

    def reset_launch_config(self):
        self._check(self._L.fg_set_launch_config(self.handle, None))

    def set_launch_table(self, entries=None):
        """Installs a tuned launch table: a list of (LaunchKey, config dict);
        [] empties it; None restores the library's built-in table."""
        if entries is None:
            self._check(self._L.fg_set_launch_table(self.handle, None, -1))
            return
        arr = (LaunchEntry * max(1, len(entries)))()
        for i, (key, cfg) in enumerate(entries):
            arr[i].key = key
            for k, v in cfg.items():
                setattr(arr[i].cfg, k, int(v))
        self._check(self._L.fg_set_launch_table(self.handle, arr, len(entries)))

    def launch_table(self) -> list:
        """The context's table as (LaunchKey, config dict) pairs."""
        n = self._L.fg_get_launch_table(self.handle, None, 0)
        arr = (LaunchEntry * max(1, n))()
        self._L.fg_get_launch_table(self.handle, arr, n)
        return [(arr[i].key, {f: getattr(arr[i].cfg, f) for f, _ in LaunchConfig._fields_}) for i in range(n)]

    def launch_config(self, desc: PlaneDesc, algo: str) -> dict:
        cfg = LaunchConfig()
        self._check(self._L.fg_get_launch_config(self.handle, C.byref(desc), 1 if algo == "grain" else 2,
                                                 C.byref(cfg)))
        return {name: getattr(cfg, name) for name, _ in LaunchConfig._fields_}


class Batch:
    """A planned batch of pixel-wise planes on one context (fg_batch_*).
    planes: list of (PlaneDesc, lambda tensor, offsets_input tensor, out tensor)
    on the context's device; the tensors must outlive the batch."""

    def __init__(self, ctx: Context, planes, capture_graph: bool = True, stream=None):
        self._L = ctx._L
        self.ctx = ctx
        n = len(planes)
        descs = (PlaneDesc * n)(*[p[0] for p in planes])
        vp = C.c_void_p
        lams = (vp * n)(*[p[1].data_ptr() for p in planes])
        offs = (vp * n)(*[p[2].data_ptr() for p in planes])
        outs = (vp * n)(*[p[3].data_ptr() for p in planes])
        h = C.c_void_p()
        rc = self._L.fg_batch_create(ctx.handle, descs, n, lams, offs, outs, 1 if capture_graph else 0,
                                     C.c_void_p(_stream_handle(stream)), C.byref(h))
        ctx._check(rc)
        self.handle = h
        self._keep = planes

    def run(self, stream=None):
        self.ctx._check(self._L.fg_batch_run(self.handle, C.c_void_p(_stream_handle(stream))))

    def status(self, stream=None):
        self.ctx._check(self._L.fg_batch_status(self.handle, C.c_void_p(_stream_handle(stream))))

    def info(self) -> dict:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._L.fg_batch_info(self.handle, C.byref(a), C.byref(b), C.byref(c))
        return dict(items=a.value, grid=b.value, graph=bool(c.value))

    def close(self):
        if self.handle:
            self._L.fg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_handle(stream) -> int:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def launch_bucket(desc: PlaneDesc, algo: str) -> LaunchKey:
    """The tuned launch table's bucket of a plane (fg_launch_bucket; no
    device needed)."""
    key = LaunchKey()
    rc = lib().fg_launch_bucket(C.byref(desc), 1 if algo == "grain" else 0, C.byref(key))
    if rc != FG_OK:
        raise FilmGrainError(rc, "fg_launch_bucket failed")
    return key


def bucket_name(key: LaunchKey) -> str:
    """'pixel|lognorm|zq4|rh4|N256': the launch table's JSON key."""
    return (f"{'grain' if key.algo == 1 else 'pixel'}|{'lognorm' if key.dist_kind else 'const'}|"
            f"zq{key.zoom_q}|rh{key.ratio_h}|N{key.n_bucket}")
