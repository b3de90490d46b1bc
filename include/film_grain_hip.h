/*
 * film_grain_hip.h -- C ABI of the MI355X (gfx950) film-grain backend.
 *
 * This is the drop-in boundary for the reference's per-plane integrator
 * dispatch (src/lib.rs:151-165): the two CPU integrators
 *     render_pixelwise(lambda, params, derived, cancel) -> Plane   (src/pixelwise.rs:11-16)
 *     render_grainwise(lambda, params, derived, cancel) -> Plane   (src/grainwise.rs:12-17)
 * and the GPU backend they sit beside
 *     wgpu::context() -> Arc<GpuContext>                           (src/wgpu/mod.rs:84)
 *     render_pixelwise_gpu(ctx, lambda, params, derived) -> Plane  (src/wgpu/mod.rs:336-341)
 *     render_grainwise_gpu(ctx, lambda, params, derived) -> Plane  (src/wgpu/mod.rs:473-478)
 * A Rust FFI crate (or ctypes, cgo, JNI ...) binds these declarations
 * unchanged: plain pointers and sizes, POD structs, int status codes.
 *
 * Results equal the reference's CPU path on the same seed (not its wgpu path,
 * whose random streams differ): the same cell/pixel hashes, Xoshiro256++
 * streams, Poisson counts and grain geometry, and the same window and disk
 * tests in f32 without contraction.
 *
 * Data contract (src/wgpu/mod.rs:17-38, 353-358, 490-495):
 *   lambda   in_h x in_w f32, row-major: the lambda_plane of one colour plane.
 *            Accepted values are those lambda_plane produces, at most
 *            MAX_LAMBDA = 1e6 (src/model.rs:11, 262).  Values <= 0 (and -inf)
 *            are skipped, as the reference skips them (src/pixelwise.rs:73,
 *            src/grainwise.rs:41).  NaN (on which the reference's Poisson::new
 *            fails), +inf and values above 1e6 are skipped too, and the call
 *            reports FG_EINVAL (at once for the host-memory entry points, from
 *            fg_stream_status for the device ones); no kernel draws on them.
 *   offsets  n_samples (x, y) f32 pairs: Derived::offsets_input (input pixels)
 *            for the pixel-wise path, Derived::offsets (output pixels) for the
 *            grain-wise path
 *   out      out_h x out_w f32, written completely
 */
#ifndef FILM_GRAIN_HIP_H
#define FILM_GRAIN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; each maps to a RenderError variant (src/lib.rs:28-44). */
#define FG_OK 0
#define FG_EINVAL 1         /* bad argument, e.g. offsets count != N (src/wgpu/mod.rs:353-358) -> Gpu(..) */
#define FG_ENODEV 2         /* no HIP device ("no compatible GPU adapter found", src/wgpu/mod.rs:101) */
#define FG_ENOMEM 3         /* device allocation failed (fatal; the context's buffers are dropped,
                               src/wgpu/mod.rs:727-736) */
#define FG_ECANCELLED 4     /* the cancel callback returned non-zero -> Cancelled */
#define FG_EHIP 5           /* a HIP runtime call or launch failed -> Gpu(..) */
#define FG_ENEIGHBOURHOOD 6 /* one output pixel's cell window does not fit in LDS (cell size
                               far below the grain radius) -> Gpu(..) */

/* Opaque per-device context (~ GpuContext, src/wgpu/mod.rs:40-49).  All entry
 * points taking a context are thread-safe (an internal mutex guards the cached
 * device buffers). */
typedef struct fg_ctx fg_ctx;

/* One colour plane's render description (~ Uniforms, src/wgpu/mod.rs:17-38,
 * built like build_uniforms :661-692, but the seed keeps all 64 bits as
 * Params::seed does on the CPU path). */
typedef struct {
    uint32_t in_w, in_h;     /* lambda plane size */
    uint32_t out_w, out_h;   /* Derived::output_width/height */
    uint32_t n_samples;      /* Params::n_samples (>= 1) */
    uint32_t dist_kind;      /* 0 = RadiusDist::Const, 1 = RadiusDist::Lognorm */
    uint64_t seed;           /* Params::seed */
    float zoom;              /* Params::zoom */
    float delta;             /* Derived::delta (pixel-wise cell size) */
    float rm;                /* Derived::rm (radius clamp) */
    float radius_mean;       /* Params::radius_mean (the Const radius) */
    float log_mu;            /* Params::radius_log_mu (f32; widened on device like RadiusProfile::new) */
    float log_sigma;         /* Params::radius_log_sigma */
} fg_plane_desc;

/* Cancellation poll (~ the cancel closure, src/lib.rs:114-123); non-zero cancels. */
typedef int (*fg_cancel_fn)(void *user);

/* Creates the context for HIP device `device_ordinal` (~ wgpu::context() / init,
 * src/wgpu/mod.rs:84-118).  FG_ENODEV when no device exists. */
int fg_ctx_create(int device_ordinal, fg_ctx **out);
void fg_ctx_destroy(fg_ctx *ctx);
/* Text of the last error on this context (~ RenderError::Gpu(String)). */
const char *fg_last_error(const fg_ctx *ctx);

/* Pixel-wise render of one plane from host memory; synchronous
 * (~ render_pixelwise_gpu, src/wgpu/mod.rs:336-471; CPU twin render_pixelwise,
 * src/pixelwise.rs:11-45).  `cancel` (may be NULL) is polled before the launch,
 * as the reference's GPU arm does (src/lib.rs:155). */
int fg_render_pixelwise(fg_ctx *ctx, const fg_plane_desc *desc, const float *lambda,
                        const float *offsets_input_xy, float *out, fg_cancel_fn cancel,
                        void *cancel_user);

/* Grain-wise render of one plane from host memory; synchronous
 * (~ render_grainwise_gpu, src/wgpu/mod.rs:473-659; CPU twin render_grainwise,
 * src/grainwise.rs:12-124). */
int fg_render_grainwise(fg_ctx *ctx, const fg_plane_desc *desc, const float *lambda,
                        const float *offsets_xy, float *out, fg_cancel_fn cancel,
                        void *cancel_user);

/* Device-resident variants: all three arrays are device pointers on the
 * context's device; the work is enqueued on `hip_stream` (a hipStream_t, NULL
 * for the default stream) and the call returns without waiting.  Errors found
 * on the device are reported by fg_stream_status.  No counterpart in the
 * reference (its GPU arm always copies through host memory); these are the
 * entry points for callers whose planes already live in HBM. */
int fg_render_pixelwise_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                               const float *d_offsets_input_xy, float *d_out, void *hip_stream);
int fg_render_grainwise_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                               const float *d_offsets_xy, float *d_out, void *hip_stream);
/* The planes of one frame on the device: n_planes (1-3) lambda planes rendered
 * with one description, as n_planes calls of fg_render_pixelwise_device give
 * them.  The reference renders a frame's planes one after another with the
 * same seed (src/lib.rs:146-153), so their cells draw the same random streams
 * (src/rng.rs:26-29) and only the Poisson thresholds differ; here every
 * cell's stream is hashed once for all planes.  Asynchronous on hip_stream. */
int fg_render_pixelwise_planes_device(fg_ctx *ctx, const fg_plane_desc *desc, int n_planes,
                                      const float *const *d_lambdas, const float *d_offsets_input_xy,
                                      float *const *d_outs, void *hip_stream);

/* Allocation-free variant of fg_render_pixelwise_device, with the same
 * arguments, result and asynchrony.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or
 * allocating library routine: every scratch buffer of the pixel-wise HBM
 * path (thresholds, bins, grains, sample table, candidate lists, tile order)
 * comes from one static device array of FG_NOALLOC_SCRATCH_BYTES (256 MiB)
 * per device, held from the library's load until the process ends.  A plane
 * whose scratch would exceed it (about 15 million grains) is refused with
 * FG_ECAPACITY; so are n_samples > 2048.  Every call on a device shares that
 * device's array, across contexts: calls that overlap on one device race
 * (serialise them, e.g. on one stream with no other caller).  The array
 * keeps the last call's contents; the path clears what must start at zero.
 * Like the other device entry points it synchronises once with the stream
 * after its counting pass. */
#define FG_NOALLOC_SCRATCH_BYTES (256u << 20)
int fg_render_pixelwise_device_noalloc(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                                       const float *d_offsets_input_xy, float *d_out, void *hip_stream);

/* Adaptive sampling: pixel-wise render where output pixel i uses only the
 * first d_budget[i] of the n_samples offsets (u16 per output pixel, clamped to
 * [1, n_samples]) and is (covered samples) * (1 / d_budget[i]): for every
 * pixel exactly what fg_render_pixelwise_device gives with n_samples =
 * d_budget[i] and the first d_budget[i] offsets (Derived::offsets_input of a
 * smaller N is a prefix of a larger N's, src/rng.rs:9-24).  Each tile does
 * the work of the smallest level (16, 32, 64, ... samples, then n_samples)
 * that covers its largest budget.  No reference counterpart; needs the HBM
 * splat formulation (n_samples <= 2048). */
int fg_render_pixelwise_adaptive_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                                        const float *d_offsets_input_xy, const uint16_t *d_budget, float *d_out,
                                        void *hip_stream);

/* Cancellation (src/pixelwise.rs:28 polls `cancel` per row; the viewer drops
 * a stale job, src/bin/viewer.rs:952-965, 1006-1010).  The context owns one
 * cancel flag in host memory that every tile of the integrators (single
 * plane, grain-wise and the batch's persistent kernel) reads before it
 * starts; a raised flag makes the remaining tiles return at once and the
 * status report FG_ECANCELLED (fg_stream_status, fg_batch_status); the
 * output is then incomplete.  fg_cancel(ctx, 1) raises it from any thread,
 * fg_cancel(ctx, 0) lowers it for the next job.  The host-memory entry
 * points (fg_render_pixelwise, fg_render_grainwise, fg_render_frame) lower
 * it on entry, poll their callback while the device works, raise it when the
 * callback asks, and lower it again on return.  No reference counterpart as
 * an entry point. */
int fg_cancel(fg_ctx *ctx, int on);

/* Waits for `hip_stream` and returns the device-side status of the work
 * enqueued by the device-resident entry points since the previous call. */
int fg_stream_status(fg_ctx *ctx, void *hip_stream);

/* Instrumented pixel-wise render (device pointers, synchronous): also returns
 * counters[0] = (pixel, sample) points, counters[1] = disk tests the reference
 * performs in its own loop order with its early exit (the T of the FP32 FLOP
 * count 16 + 6T), counters[2] = covered points.  No reference counterpart. */
int fg_pixelwise_stats_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                              const float *d_offsets_input_xy, float *d_out, uint64_t counters[3],
                              void *hip_stream);

/* Instrumented grain-wise render, counting the reference's own work
 * (src/grainwise.rs:46-104): counters[0] = grains drawn with a positive
 * output radius, counters[1] = (grain, sample) pairs = grains * N,
 * counters[2] = disk tests (pixels of a pair's bounds() box in rows with
 * dy^2 <= R^2), counters[3] = tests that pass.  No reference counterpart. */
int fg_grainwise_stats_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                              const float *d_offsets_xy, float *d_out, uint64_t counters[4],
                              void *hip_stream);

/* Launch shape used by the next renders on this context; 0 keeps the
 * automatic choice.  No reference counterpart (the autotuner sets these). */
typedef struct {
    int pix_tile_w, pix_tile_h, pix_lds_bytes;
    int grain_tile_w, grain_tile_h, grain_lds_bytes, grain_grains_per_thread;
    int pix_variant; /* pixel-wise formulation: 0 automatic (= 3), 1 gather in the reference's
                        loop order, 2 splat with per-tile grain generation, 3 splat from grains
                        generated once per plane in HBM; all give identical results */
    int pix_block;   /* threads per workgroup of the HBM splat kernel: 512 (default) or 256 */
    int grain_variant; /* grain-wise formulation: 0 automatic (= 2), 1 the reference's loop order
                          over every (grain, sample) pair, 2 runs of the sample order sorted by
                          offset fraction; identical results */
    int grain_block; /* threads per workgroup of the sorted grain-wise kernel: 512 (default) or 256 */
} fg_launch_config;
/* An explicit launch configuration (any field nonzero) replaces the tuned
 * table below for the following renders; NULL (or all zeros) returns to it. */
int fg_set_launch_config(fg_ctx *ctx, const fg_launch_config *cfg);
/* The launch shape the next render of desc with algo (0 pixel-wise, 1
 * grain-wise) takes: the explicit configuration, else the tuned table's entry
 * for the plane's bucket, else the built-in choice. */
int fg_get_launch_config(fg_ctx *ctx, const fg_plane_desc *desc, int algo, fg_launch_config *out);

/* Tuned launch table.  No reference counterpart: the reference dispatches
 * every plane with one fixed workgroup shape (src/wgpu/mod.rs:439-441).  The
 * autotuner (film_grain_amd/autotune.py) times candidate shapes per workload
 * bucket and records the fastest; every candidate gives the same result, so
 * the table changes only speed.  A context starts with the table the library
 * was built with (film_grain_amd/tuning/launch_table.json).  Buckets: the
 * integrator, the radius law, zoom in quarter octaves, the radius ratio (rm /
 * cell size for pixel-wise, rm * zoom for grain-wise) in half octaves, and N
 * rounded up to a power of two. */
typedef struct {
    uint32_t algo;      /* 0 pixel-wise, 1 grain-wise */
    uint32_t dist_kind; /* 0 constant, 1 lognormal */
    int32_t zoom_q;     /* round(4 log2 zoom) */
    int32_t ratio_h;    /* round(2 log2 ratio); -1000 for ratio <= 0 */
    uint32_t n_bucket;  /* smallest power of two >= N */
} fg_launch_key;
typedef struct {
    fg_launch_key key;
    fg_launch_config cfg;
} fg_launch_entry;
/* The bucket of desc for algo (no context or device needed). */
int fg_launch_bucket(const fg_plane_desc *desc, int algo, fg_launch_key *out);
/* Replaces the context's table with n entries (copied); n = 0 empties it,
 * n < 0 restores the built-in table.  fg_get_launch_table copies up to cap
 * entries and returns the table's size. */
int fg_set_launch_table(fg_ctx *ctx, const fg_launch_entry *entries, int n);
int fg_get_launch_table(fg_ctx *ctx, fg_launch_entry *out, int cap);

/* Kernel timing of the following renders on this context: with enable set,
 * HIP events are recorded on the launch stream around the grain-generation
 * kernels and around the coverage-test kernel of each render, and
 * fg_get_timings waits for the last render and returns milliseconds:
 * ms[0] whole render, ms[1] generation kernels, ms[2] coverage-test kernel.
 * No reference counterpart (benchmarking support). */
int fg_set_timing(fg_ctx *ctx, int enable);
int fg_get_timings(fg_ctx *ctx, float ms[3]);

/* ---- Batched rendering -------------------------------------------------
 * No counterpart in the reference, which renders one plane per call
 * (src/color.rs:47-64 loops over planes; src/lib.rs:151-165 dispatches each).
 * A batch is a fixed set of 1 to 65535 pixel-wise planes of any sizes and
 * parameters (N <= 2048 each), with their device buffers.  fg_batch_create
 * plans it: per-plane geometry and launch shapes, buffers sized from one
 * counting pass over the current inputs (grain capacity = 1.25 x that count +
 * 65536), and, with capture_graph set, one HIP graph of the whole pipeline:
 * the grain generation (thresholds, draws and sample tables each one launch
 * over all planes; counts one launch per image, the consecutive planes with
 * one description and offset range counted together), one scan over all
 * planes' bins, a heaviest-first order of all (plane, tile) items, and one
 * persistent splat launch whose workgroups take items from an atomic
 * counter.  fg_batch_run
 * enqueues the pipeline (or replays the graph) on `hip_stream`; the inputs may
 * change between runs, and a run whose grains exceed the capacity renders
 * nothing and makes fg_batch_status return FG_ECAPACITY (create the batch
 * again).  Each plane's result equals fg_render_pixelwise_device's bit for
 * bit.  A batch must not run concurrently with itself. */
#define FG_ECAPACITY 7 /* also: a plane (or batch) with more than fg_max_grains() grains */
typedef struct fg_batch fg_batch;
int fg_batch_create(fg_ctx *ctx, const fg_plane_desc *descs, int n_planes, const float *const *d_lambdas,
                    const float *const *d_offsets_input_xy, float *const *d_outs, int capture_graph,
                    void *hip_stream, fg_batch **out);
int fg_batch_run(fg_batch *batch, void *hip_stream);
/* Waits for `hip_stream` and returns the status of the batch's last run. */
int fg_batch_status(fg_batch *batch, void *hip_stream);
/* Number of (plane, tile) work items and the persistent grid size. */
int fg_batch_info(const fg_batch *batch, int *n_items, int *grid_blocks, int *graph_captured);
void fg_batch_destroy(fg_batch *batch);

/* ---- Per-pixel Monte-Carlo error model -----------------------------------
 * No counterpart in the reference.  A small MLP (16 inputs, 12 used -> 64 ->
 * 64 -> 1, ReLU) predicts a log standard deviation of an output pixel's
 * Monte-Carlo error from local statistics of the normalised intensity plane
 * (trained on the Gaussian likelihood of N = 64 minus N = 4096 residuals, it
 * is the log std of the N = 64 error, sqrt(p (1 - p) (1/64 - 1/4096)) for a
 * pixel of covered fraction p).  BF16 weights and operands, FP32
 * accumulation on the matrix cores; forward, fused backward of a squared-
 * error or likelihood loss, and an Adam step, all on device memory and
 * asynchronous on `hip_stream`.  Parameters: FG_MLP_PARAMS floats: W1
 * [16][64] input-major (rows 12..15 zero), b1 [64], W2 [64][64], b2 [64],
 * w3 [64], b3.
 * Features: FG_MLP_FEATURES BF16 values per output pixel:
 *   u, 3x3 mean, 3x3 std, gradient magnitude, u(1-u), radius/0.3,
 *   radius_stddev/0.15, sigma/2, zoom/4, channel one-hot (3), 4 zeros. */
#define FG_MLP_PARAMS 5313
#define FG_MLP_FEATURES 16
int fg_mlp_features_device(fg_ctx *ctx, const float *d_u, uint32_t in_w, uint32_t in_h, uint32_t out_w,
                           uint32_t out_h, float zoom, float radius_mean, float radius_stddev, float sigma,
                           int channel, uint16_t *d_features, void *hip_stream);
int fg_mlp_forward_device(fg_ctx *ctx, const float *d_params, const uint16_t *d_features, int64_t n, float *d_out,
                          void *hip_stream);
/* Adds the gradient of (1/n_total) sum L over these n rows to d_grad and the
 * sum of L to *d_loss.  FG_LOSS_SQUARED: L = (y - target)^2.  FG_LOSS_NLL:
 * the Gaussian negative log-likelihood of the residual target under the
 * predicted log std y, L = y + target^2 exp(-2y) / 2. */
#define FG_LOSS_SQUARED 0
#define FG_LOSS_NLL 1
int fg_mlp_grad_device(fg_ctx *ctx, const float *d_params, const uint16_t *d_features, const float *d_target,
                       int64_t n, int64_t n_total, int loss_kind, float *d_grad, float *d_loss, void *hip_stream);
/* Per-pixel sample budgets for fg_render_pixelwise_adaptive_device:
 * budget = clamp(ceil(c * s), n_min, n_max) with s = exp(d_pred_log_s[i]), or,
 * with d_pred_log_s NULL, s = sqrt(u (1 - u)) from the features (the
 * allocator without the model). */
int fg_mlp_budget_device(fg_ctx *ctx, const float *d_pred_log_s, const uint16_t *d_features, int64_t n, float c,
                         int n_min, int n_max, uint16_t *d_budget, void *hip_stream);
/* The same budgets with the scale read from device memory (*d_c, as
 * fg_budget_scale_device writes it at d_scale + 1), so that a frame's scale
 * search and budgets run without a host round trip.  No counterpart in the
 * reference (its sample count is one N for every pixel). */
int fg_mlp_budget_scale_device(fg_ctx *ctx, const float *d_pred_log_s, const uint16_t *d_features, int64_t n,
                               const float *d_c, int n_min, int n_max, uint16_t *d_budget, void *hip_stream);
/* The scale of a frame's budgets on the device: the smallest c whose budgets
 * clamp(ceil(c s), n_min, n_max), s = exp(pred) * s_mult, have mean predicted
 * squared error s^2 (1/n - 1/n_ref) <= target_mse, over a strided subsample
 * of at most 2^18 of the n rows, by three rounds of 64 geometric candidates.
 * Writes d_scale[0] = c and d_scale[1] = c * s_mult.  No counterpart in the
 * reference. */
int fg_budget_scale_device(fg_ctx *ctx, const float *d_pred_log_s, int64_t n, float s_mult, double target_mse,
                           int n_min, int n_max, int n_ref, float *d_scale, void *hip_stream);
/* One Adam step (step counts from 1) on d_params with moments d_m, d_v. */
int fg_mlp_adam_device(fg_ctx *ctx, float *d_params, const float *d_grad, float *d_m, float *d_v, float lr,
                       float beta1, float beta2, float eps, int step, void *hip_stream);

/* ---- Frames: colour split, lambda pre-pass, render, merge on the device ---
 * The reference prepares each plane on the host (src/color.rs:158-213 split,
 * src/model.rs:228-265 normalize_plane and lambda_plane), renders it, and
 * merges and quantises on the host (src/color.rs:66-114, 237-239).  These
 * entry points do the same on the device, bit for bit: the same f32
 * operations in the same order, and the C library's logf (which the
 * reference's f32::ln calls) reproduced exactly on [1e-6, 1]. */
#define FG_COLOR_LUMA 0 /* ColorMode::Luma: render Y, carry BT.709 Cb/Cr (resized nearest) */
#define FG_COLOR_RGB 1  /* ColorMode::Rgb: render R, G, B */
#define FG_ALGO_PIXEL 0
#define FG_ALGO_GRAIN 1
typedef struct {
    fg_plane_desc plane;  /* shared by the frame's planes; in_w/in_h = the frame (after any ROI crop) */
    uint32_t color_mode;  /* FG_COLOR_LUMA or FG_COLOR_RGB */
    uint32_t algo;        /* FG_ALGO_PIXEL or FG_ALGO_GRAIN (choose_algorithm's result) */
    float inv_e_pi_r2;    /* Derived::inv_e_pi_r2 */
    uint32_t reserved;    /* 0 */
} fg_frame_desc;
/* One frame: d_rgb is in_h x in_w interleaved RGB f32 (as image's to_rgb32f
 * gives); d_offsets as for the plane entry points of fd->algo; d_rgb8 gets
 * out_h x out_w interleaved 8-bit RGB (into_rgb_image); d_planes, if not
 * NULL, gets the rendered planes (3 for RGB, 1 for luma), out_h x out_w f32
 * each.  Asynchronous on hip_stream. */
int fg_render_frame_device(fg_ctx *ctx, const fg_frame_desc *fd, const float *d_rgb, const float *d_offsets_xy,
                           uint8_t *d_rgb8, float *d_planes, void *hip_stream);
/* The same from and to host memory: one upload of the frame and the offsets,
 * one download of the result; cancel is polled between planes.  Synchronous. */
int fg_render_frame(fg_ctx *ctx, const fg_frame_desc *fd, const float *rgb, const float *offsets_xy, uint8_t *rgb8,
                    float *planes, fg_cancel_fn cancel, void *cancel_user);
/* The steps on their own (device memory, asynchronous). */
int fg_lambda_device(fg_ctx *ctx, const float *d_plane, uint32_t w, uint32_t h, float inv_e_pi_r2, float *d_norm,
                     float *d_lambda, void *hip_stream); /* normalize_plane + lambda_plane; d_norm may be NULL */
int fg_color_split_device(fg_ctx *ctx, int color_mode, const float *d_rgb, uint32_t w, uint32_t h, float *d_p0,
                          float *d_p1, float *d_p2, void *hip_stream);
/* Luma: d_p0 = Y at w x h, d_p1/d_p2 = Cb/Cr at cw x ch; RGB: all w x h. */
int fg_color_merge_device(fg_ctx *ctx, int color_mode, const float *d_p0, const float *d_p1, const float *d_p2,
                          uint32_t w, uint32_t h, uint32_t cw, uint32_t ch, uint8_t *d_rgb8, void *hip_stream);
/* d_out[i] = logf(d_in[i]) exactly as the C library computes it, for inputs in
 * [1e-6, 1] (other inputs get a result within one unit in the last place). */
int fg_log_f32_device(fg_ctx *ctx, const float *d_in, int64_t n, float *d_out, void *hip_stream);

/* ---- Generator inspection (parity tests) --------------------------------
 * No reference counterpart: they expose the device generators' state so that
 * tests can compare it cell by cell with the reference's CPU path.
 *
 * fg_pixelwise_grains runs the pixel-wise HBM path's generation for one plane
 * (per cell: cell_rng, the clamped lambda lookup, the Poisson count, then each
 * grain's centre and radius, src/pixelwise.rs:66-100) and copies every grain
 * record to host memory as 4 floats (cx, cy, r, cell), where cell holds the
 * bits (i - origin[0]) | (j - origin[1]) << 16 of the grain's cell (i, j);
 * each cell's grains are contiguous and in draw order.  origin and size get
 * the plane's extended cell grid (first cell, cells per row and column);
 * *n_grains the plane's grain count.  When that exceeds capacity nothing is
 * copied and FG_ECAPACITY is returned.  Synchronous.
 *
 * fg_grainwise_pixel_state runs the grain-wise path's Poisson pass for one
 * plane (pixel_rng and Poisson(lambda), src/grainwise.rs:37-45) and copies,
 * per input pixel, the count to q_out[in_w * in_h] and the Xoshiro256++
 * state after the count draws (from which the grain draws continue) to
 * state_out[4 * in_w * in_h].  Synchronous. */
int fg_pixelwise_grains(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                        const float *d_offsets_input_xy, float *grains_xyrc, uint64_t capacity,
                        uint64_t *n_grains, int32_t origin[2], uint32_t size[2]);
int fg_grainwise_pixel_state(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda, uint32_t *q_out,
                             uint64_t *state_out);

/* Most grains one pixel-wise plane (or one batch) may hold: grain indices,
 * bin offsets and chunk indices are 32-bit.  The count pass sums the grains in
 * 64 bits, and a plane or batch above this limit is refused with
 * FG_ECAPACITY before any grain is written.  No reference counterpart (the CPU
 * path never stores grains). */
uint64_t fg_max_grains(void);

/* The f64 math the integrators evaluate on the device, element-wise on n
 * doubles (device memory, asynchronous), for comparison with the C library
 * the reference's Rust std and rand_distr call: FG_MATH_EXP_NEG exp(-x) (the
 * Knuth threshold, src/pixelwise.rs:81 via Poisson::new), FG_MATH_EXP exp(x),
 * FG_MATH_LOG log(x), FG_MATH_TAN_PI tan(pi * x) (the Cauchy draw of the
 * rejection sampler), FG_MATH_LOG_GAMMA rand_distr's log_gamma(x),
 * FG_MATH_SQRT sqrt(x), FG_MATH_EXP_F32 (double)(float)exp(x) (a LogNormal
 * radius, src/model.rs:142).  No reference counterpart. */
#define FG_MATH_EXP_NEG 0
#define FG_MATH_EXP 1
#define FG_MATH_LOG 2
#define FG_MATH_TAN_PI 3
#define FG_MATH_LOG_GAMMA 4
#define FG_MATH_SQRT 5
#define FG_MATH_EXP_F32 6
int fg_math_device(fg_ctx *ctx, int fn, const double *d_in, int64_t n, double *d_out, void *hip_stream);

/* Library version string. */
const char *fg_version(void);

#ifdef __cplusplus
}
#endif
#endif
