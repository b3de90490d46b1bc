// fg_capi.hip -- C ABI of the gfx950 film-grain backend (include/film_grain_hip.h).
//
// Host side of the boundary that replaces the reference's per-plane GPU arm
// (src/lib.rs:151-165, src/wgpu/mod.rs:336-659): argument checks, a cached
// per-device context, launch-shape choice, and the synchronous host-memory
// entry points built on the asynchronous device-memory ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/film_grain_hip.h"
#include "fg_frame.h"
#include "fg_kernels.h"
#include "fg_mlp.h"
#include "fg_pixelwise_hbm.h"
#include "fg_rng.h"
#include "fg_launch_table.h"

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct fg_ctx {
    int device = 0;
    int max_lds = 0;      // per-workgroup LDS limit
    int num_cus = 0;
    std::mutex mu;
    std::string last_error;
    hipStream_t stream = nullptr;  // stream of the host-memory entry points
    // cached device buffers of the host-memory entry points
    float *d_lambda = nullptr;
    size_t cap_lambda = 0;
    float *d_offs = nullptr;
    size_t cap_offs = 0;
    float *d_out = nullptr;
    size_t cap_out = 0;
    // per-context device scratch: status word, stats counters, offsets range
    int *d_status = nullptr;
    unsigned long long *d_stats = nullptr;
    unsigned long long *d_sum64 = nullptr;  // a plane's 64-bit grain total (bin_count)
    // device address of this device's copy of the allocation-free entry
    // point's static scratch array (hipGetSymbolAddress at creation)
    unsigned char *noalloc_scratch = nullptr;
    // cancel flag: coherent host memory that every tile of the integrators
    // reads once (host and device addresses)
    int *h_cancel = nullptr;
    int *d_cancel = nullptr;
    float4 *d_range = nullptr;
    // scratch of the per-plane grain generation: Poisson thresholds, bin
    // counts and starts, scan workspace, grain records
    DevBuf thresholds, bin_count, bin_start, scan_temp, grains;
    DevBuf samples;   // sample order of the splat kernel
    DevBuf bin_mask;  // nonempty cells per bin (256 bits)
    DevBuf plane_counts;  // planes of one frame: thresholds, bin counts and masks per plane
    DevBuf pix_state; // grain-wise: per input pixel grain count and stream state (36 B)
    DevBuf tiles, tile_temp;  // heaviest-first tile order: weights and ids (in, out), sort workspace
    DevBuf levels;            // adaptive sampling: level sizes, level tables, per-tile levels
    DevBuf frame;             // frames: split planes, lambda, max word, rendered planes
    DevBuf cells;             // candidate lists: grain-wise constant radius, splat direct test
    DevBuf scale_state;       // budget scale search: bracket, flag, probe sums
    DevBuf frame_io;          // frames from host memory: RGB in, offsets, 8-bit RGB out, planes out
    std::vector<int> level_n;
    std::vector<fg::SampleTable> level_tables;

    hipEvent_t scratch_done = nullptr;  // recorded after the last kernel that reads the scratch
    bool scratch_pending = false;
    // optional kernel timing: start, generation done, test kernel start, end
    bool timing = false;
    bool timed = false;
    hipEvent_t tev[4] = {nullptr, nullptr, nullptr, nullptr};
    fg_launch_config cfg{};
    bool cfg_explicit = false;            // cfg set by the caller: it replaces the table
    std::vector<fg_launch_entry> table;   // tuned launch shapes per workload bucket
    std::map<uint32_t, bool> div_checked;  // cell size -> reciprocal division verified
};

// Scratch of fg_render_pixelwise_device_noalloc: one copy per device, held
// from the library's load until the process ends.  In host code this name is
// a host-side stand-in, never a device address: host code uses the address
// hipGetSymbolAddress returns for the current device (fg_ctx::noalloc_scratch).
constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Most grains one plane (or batch) may hold: grain indices, bin starts and
// the splat's chunk indices (up to 63 past the last grain) are u32.
constexpr unsigned long long kMaxGrains = 0xFFFFFF00ull;

int set_error(fg_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->last_error = msg;
    return code;
}

int hip_error(fg_ctx *ctx, hipError_t e, const char *what)
{
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        return set_error(ctx, FG_ENOMEM,
                         std::string(what) +
                             " ran out of GPU memory. Try lowering the output size or sample count.");
    }
    return set_error(ctx, FG_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void free_buf(DevBuf &b)
{
    (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

void free_buffers(fg_ctx *ctx)
{
    (void)hipFree(ctx->d_lambda);
    (void)hipFree(ctx->d_offs);
    (void)hipFree(ctx->d_out);
    ctx->d_lambda = ctx->d_offs = ctx->d_out = nullptr;
    ctx->cap_lambda = ctx->cap_offs = ctx->cap_out = 0;
    free_buf(ctx->thresholds);
    free_buf(ctx->bin_count);
    free_buf(ctx->bin_start);
    free_buf(ctx->scan_temp);
    free_buf(ctx->grains);
    free_buf(ctx->samples);
    free_buf(ctx->bin_mask);
    free_buf(ctx->plane_counts);
    free_buf(ctx->pix_state);
    free_buf(ctx->tiles);
    free_buf(ctx->tile_temp);
    free_buf(ctx->levels);
    free_buf(ctx->frame);
    free_buf(ctx->frame_io);
    free_buf(ctx->cells);
    free_buf(ctx->scale_state);
}

int ensure_bytes(fg_ctx *ctx, DevBuf &b, size_t bytes, const char *what)
{
    if (b.cap >= bytes) return FG_OK;
    free_buf(b);
    const size_t want = std::max<size_t>(bytes + bytes / 8, 256);  // headroom for the next plane
    const hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        free_buffers(ctx);
        (void)hipGetLastError();
        return hip_error(ctx, e, what);
    }
    b.cap = want;
    return FG_OK;
}

// The generation scratch is shared by every call on the context; a call waits
// until the previous user's last kernel has finished before overwriting it.
int scratch_acquire(fg_ctx *ctx, const char *what)
{
    if (!ctx->scratch_pending) return FG_OK;
    const hipError_t e = hipEventSynchronize(ctx->scratch_done);
    ctx->scratch_pending = false;
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, what);
}

void mark(fg_ctx *ctx, int i, hipStream_t s)
{
    if (!ctx->timing) return;
    if (!ctx->tev[i]) (void)hipEventCreate(&ctx->tev[i]);
    (void)hipEventRecord(ctx->tev[i], s);
    if (i == 3) ctx->timed = true;
}

int scratch_release(fg_ctx *ctx, hipStream_t s, const char *what)
{
    const hipError_t e = hipEventRecord(ctx->scratch_done, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    ctx->scratch_pending = true;
    return FG_OK;
}

int ensure(fg_ctx *ctx, float **buf, size_t *cap, size_t n_floats, const char *what)
{
    if (*cap >= n_floats) return FG_OK;
    (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    const size_t bytes = std::max<size_t>(n_floats, 1) * sizeof(float);
    const hipError_t e = hipMalloc(buf, bytes);
    if (e != hipSuccess) {
        // An allocation failure is fatal for the cached state, as the reference
        // drops its context after an out-of-memory error.
        free_buffers(ctx);
        (void)hipGetLastError();
        return hip_error(ctx, e, what);
    }
    *cap = n_floats;
    return FG_OK;
}

int validate(fg_ctx *ctx, const fg_plane_desc *d)
{
    if (!ctx) return FG_EINVAL;
    if (!d) return set_error(ctx, FG_EINVAL, "null plane description");
    if (d->in_w == 0 || d->in_h == 0) return set_error(ctx, FG_EINVAL, "input image is empty");
    if (d->out_w == 0 || d->out_h == 0)
        return set_error(ctx, FG_EINVAL, "output dimensions must be positive");
    if (d->in_w > (1u << 24) || d->in_h > (1u << 24) || d->out_w > (1u << 24) ||
        d->out_h > (1u << 24))
        return set_error(ctx, FG_EINVAL, "image dimensions exceed 2^24");
    if (d->n_samples == 0) return set_error(ctx, FG_EINVAL, "n_samples must be >= 1");
    if (d->dist_kind > 1) return set_error(ctx, FG_EINVAL, "unknown radius distribution");
    if (!(d->zoom > 0.0f) || !std::isfinite(d->zoom)) return set_error(ctx, FG_EINVAL, "zoom must be > 0");
    if (!(d->delta > 0.0f) || !std::isfinite(d->delta)) return set_error(ctx, FG_EINVAL, "cell size must be > 0");
    if (!(d->rm > 0.0f) || !std::isfinite(d->rm)) return set_error(ctx, FG_EINVAL, "max radius must be > 0");
    if (!std::isfinite(d->radius_mean)) return set_error(ctx, FG_EINVAL, "radius must be finite");
    if (d->dist_kind == 1 && (!std::isfinite(d->log_mu) || !(d->log_sigma >= 0.0f) ||
                              !std::isfinite(d->log_sigma)))
        return set_error(ctx, FG_EINVAL, "invalid log-normal parameters");
    return FG_OK;
}

// Gather kernel: largest tile (from a fixed ladder) whose cell neighbourhood
// index fits in about half of the LDS budget, for an assumed offset spread of
// +-3.5 output px.  Tiles whose actual neighbourhood is larger are split inside
// the kernel, so this is a speed choice, not a correctness one.
void choose_gather_tile(const fg_plane_desc *d, int lds, int *tw, int *th)
{
    static const int ladder[][2] = {{32, 32}, {32, 16}, {16, 16}, {16, 8}, {8, 8}, {8, 4}, {4, 4}};
    const double iz = 1.0 / d->zoom;
    const double spread = 2.0 * 3.5 * iz;
    for (const auto &t : ladder) {
        const double W = ((t[0] - 1) * iz + spread + 2.0 * d->rm) / d->delta + 3.0;
        const double H = ((t[1] - 1) * iz + spread + 2.0 * d->rm) / d->delta + 3.0;
        if ((W * H + 1) * 4.0 <= 0.5 * lds) {
            *tw = t[0];
            *th = t[1];
            return;
        }
    }
    *tw = 4;
    *th = 4;
}

// Splat kernel: largest tile whose N-bit sample sets take at most half of the
// LDS budget, leaving room for the offsets table and at least 2048 grains.
void choose_splat_tile(const fg_plane_desc *d, int lds, int *tw, int *th)
{
    static const int ladder[][2] = {{64, 64}, {64, 32}, {32, 32}, {32, 16}, {16, 16}, {16, 8},
                                    {8, 8},   {8, 4},   {4, 4},   {4, 2},   {2, 2},   {2, 1}, {1, 1}};
    const size_t words = (d->n_samples + 31) / 32;
    for (const auto &t : ladder) {
        const size_t bits = (size_t)t[0] * t[1] * words * 4;
        if (bits <= (size_t)lds / 2 && bits + 1024 * 32 <= (size_t)lds) {
            *tw = t[0];
            *th = t[1];
            return;
        }
    }
    *tw = 1;
    *th = 1;
}

// Pixel-wise formulation: 1 gather (the reference's loop order), 2 splat with
// in-tile generation, 3 splat from grains generated once per plane in HBM.
// The automatic choice (0) is 3; the caller falls back to 2 when the plane's
// cell grid is too large to pack cell indices in 16 bits.
// Largest sample count whose sorted sample table the HBM splat kernel keeps
// in LDS (18 bytes per sample plus the bucket index).
constexpr uint32_t kHbmMaxSamples = 2048;

// The workload bucket of a plane for the tuned launch table.
fg_launch_key launch_key(const fg_plane_desc *d, int algo)
{
    fg_launch_key k{};
    k.algo = algo == 1 ? 1u : 0u;
    k.dist_kind = d->dist_kind ? 1u : 0u;
    const double zoom = (double)d->zoom;
    k.zoom_q = zoom > 0.0 ? (int32_t)std::floor(4.0 * std::log2(zoom) + 0.5) : -1000;
    const double ratio = algo == 1 ? (double)d->rm * zoom : (d->delta > 0.0f ? (double)d->rm / (double)d->delta : 0.0);
    k.ratio_h = ratio > 0.0 ? (int32_t)std::floor(2.0 * std::log2(ratio) + 0.5) : -1000;
    uint32_t nb = 1;
    while (nb < d->n_samples && nb < 0x80000000u) nb <<= 1;
    k.n_bucket = nb;
    return k;
}

// The launch shape of a render: the caller's explicit configuration, else the
// tuned table's entry for the plane's bucket, else all zeros (the built-in
// choices below).
const fg_launch_config &launch_cfg(const fg_ctx *ctx, const fg_plane_desc *d, int algo)
{
    static const fg_launch_config none{};
    if (ctx->cfg_explicit) return ctx->cfg;
    const fg_launch_key k = launch_key(d, algo);
    for (const fg_launch_entry &e : ctx->table)
        if (e.key.algo == k.algo && e.key.dist_kind == k.dist_kind && e.key.zoom_q == k.zoom_q &&
            e.key.ratio_h == k.ratio_h && e.key.n_bucket == k.n_bucket)
            return e.cfg;
    return none;
}

int pixel_variant(fg_ctx *ctx, const fg_plane_desc *d)
{
    const fg_launch_config &cfg = launch_cfg(ctx, d, 0);
    if (cfg.pix_variant == 1 || d->n_samples > 65535) return 1;
    if (cfg.pix_variant == 2 || d->n_samples > kHbmMaxSamples) return 2;
    return 3;
}

int splat_hbm_block(const fg_ctx *ctx, const fg_plane_desc *d)
{
    return launch_cfg(ctx, d, 0).pix_block == 256 ? 256 : 512;
}

// Buckets of the sample order's fraction index: the next power of two >= N,
// within [64, 4096].
int sample_buckets(uint32_t n)
{
    int q = 64;
    while (q < (int)n && q < 4096) q <<= 1;
    return q;
}

// HBM splat kernel (block threads per workgroup, 0: the configured size): the
// configured tile, else a tile of the ladder chosen
// below (with min_tiles > 0, stepped down until the plane has that many
// tiles); the LDS is exactly what the tile and the sample table need.
// Returns false when no tile fits the workgroup LDS limit.  Queue entries
// hold a tile column in 7 bits and a row in 8 bits, so tiles are at most
// 128 x 256.
bool hbm_config(fg_ctx *ctx, const fg_plane_desc *d, int *tw, int *th, int *lds, long long min_tiles, int block = 0)
{
    static const int ladder[][2] = {{64, 32}, {32, 32}, {32, 16}, {16, 16}, {16, 8}, {8, 8},
                                    {8, 4},   {4, 4},   {4, 2},   {2, 2},   {2, 1}, {1, 1}};
    const int q = sample_buckets(d->n_samples);
    const size_t words = (d->n_samples + 31) / 32;
    const fg_launch_config &cfg = launch_cfg(ctx, d, 0);
    const size_t queue = fg::splat_hbm_static_lds(block > 0 ? block : splat_hbm_block(ctx, d));
    const size_t limit = (size_t)ctx->max_lds - queue;
    auto fits = [&](int w, int h) { return fg::splat_hbm_lds(w, h, (int)d->n_samples, q) <= limit; };
    // room for two workgroups per CU (at N = 2048 one workgroup per CU made
    // the splat 2x slower)
    auto fits2 = [&](int w, int h) {
        return fg::splat_hbm_lds(w, h, (int)d->n_samples, q) + queue <= (size_t)ctx->max_lds / 2;
    };
    if (cfg.pix_tile_w > 0 && cfg.pix_tile_h > 0 && cfg.pix_tile_w <= 128 &&
        cfg.pix_tile_h <= 256 && fits(cfg.pix_tile_w, cfg.pix_tile_h)) {
        *tw = cfg.pix_tile_w;
        *th = cfg.pix_tile_h;
    } else {
        // The largest ladder tile of 128 to max(128, 512 zoom^2) output
        // pixels (about 512 input pixels: larger tiles leave too few
        // workgroups to balance) whose bitset takes at most 16 KB; else one
        // with up to 32 KB; else a smaller tile with at most 16 KB; each
        // leaving room for two workgroups per CU; else any that fits
        // (measured on MI355X, 1080p zoom-1 planes, N = 16..2048).
        const double cap_px = std::max(128.0, 512.0 * (double)d->zoom * (double)d->zoom);
        *tw = 0;
        for (int pass = 0; pass < 4 && *tw == 0; ++pass) {
            for (const auto &t : ladder) {
                const int px = t[0] * t[1];
                const size_t bits = (size_t)px * words * 4;
                const bool in_cap = (double)px <= cap_px;
                const bool ok = pass == 0   ? in_cap && px >= 128 && bits <= 16384 && fits2(t[0], t[1])
                                : pass == 1 ? in_cap && px >= 128 && bits <= 32768 && fits2(t[0], t[1])
                                : pass == 2 ? in_cap && bits <= 16384 && fits2(t[0], t[1])
                                            : true;
                if (ok && fits(t[0], t[1])) {
                    *tw = t[0];
                    *th = t[1];
                    break;
                }
            }
        }
        if (*tw == 0) return false;
    }
    // A plane alone on the device: smaller tiles until it has at least four
    // per CU, so the heaviest tile does not set the frame time (480x272,
    // N = 64: 255 tiles of 32x16 take 2.6 ms).  A tuned tile steps down too
    // (its bucket holds planes of every size); an explicit one does not.
    if (min_tiles > 0 && !ctx->cfg_explicit) {
        size_t i = 0;
        while (i < sizeof ladder / sizeof ladder[0] && (ladder[i][0] != *tw || ladder[i][1] != *th)) ++i;
        auto tiles = [&](int w, int h) {
            return (long long)((d->out_w + w - 1) / w) * (long long)((d->out_h + h - 1) / h);
        };
        while (i + 1 < sizeof ladder / sizeof ladder[0] && tiles(*tw, *th) < min_tiles &&
               ladder[i + 1][0] * ladder[i + 1][1] >= 32 && fits(ladder[i + 1][0], ladder[i + 1][1])) {
            ++i;
            *tw = ladder[i][0];
            *th = ladder[i][1];
        }
    }
    *lds = (int)fg::splat_hbm_lds(*tw, *th, (int)d->n_samples, q);
    return true;
}

void pixel_config(fg_ctx *ctx, const fg_plane_desc *d, int variant, int *tw, int *th, int *lds)
{
    const fg_launch_config &cfg = launch_cfg(ctx, d, 0);
    const int def = variant == 1 ? 64 * 1024 : 72 * 1024;
    const int budget = cfg.pix_lds_bytes > 0 ? cfg.pix_lds_bytes : def;
    *lds = std::min(budget, ctx->max_lds - 6 * 1024);
    const int max_px = variant == 1 ? fg::kMaxTilePixels : 256 * 256;
    if (cfg.pix_tile_w > 0 && cfg.pix_tile_h > 0 &&
        cfg.pix_tile_w * cfg.pix_tile_h <= max_px && cfg.pix_tile_w <= 256 &&
        cfg.pix_tile_h <= 256) {
        *tw = cfg.pix_tile_w;
        *th = cfg.pix_tile_h;
    } else if (variant == 1) {
        choose_gather_tile(d, *lds, tw, th);
    } else {
        choose_splat_tile(d, *lds, tw, th);
    }
}

// floor(a / delta) through the reciprocal sequence is used only for cell
// sizes where the host has checked it against IEEE division.
bool division_fast_ok(fg_ctx *ctx, float delta, float rdelta)
{
    uint32_t key;
    std::memcpy(&key, &delta, 4);
    auto it = ctx->div_checked.find(key);
    if (it != ctx->div_checked.end()) return it->second;
    const bool ok = fg::reciprocal_division_exact(delta, rdelta);
    ctx->div_checked[key] = ok;
    return ok;
}

void grain_config(fg_ctx *ctx, const fg_plane_desc *d, int *tw, int *th, int *lds, int *gpt)
{
    const fg_launch_config &cfg = launch_cfg(ctx, d, 1);
    const int words = (int)((d->n_samples + 31) / 32);
    *gpt = cfg.grain_grains_per_thread > 0 ? cfg.grain_grains_per_thread : 4;
    const int budget = cfg.grain_lds_bytes > 0 ? cfg.grain_lds_bytes : 64 * 1024;
    const int cap = std::min(budget, ctx->max_lds - 2 * 1024);
    const int slots = fg::kBlock * (*gpt) * 16;
    if (cfg.grain_tile_w > 0 && cfg.grain_tile_h > 0) {
        *tw = cfg.grain_tile_w;
        *th = cfg.grain_tile_h;
    } else {
        static const int ladder[][2] = {{128, 64}, {64, 64}, {64, 32}, {32, 32}, {32, 16},
                                        {16, 16}, {16, 8},   {8, 8},   {8, 4},   {4, 4},
                                        {2, 2},   {1, 1}};
        *tw = 1;
        *th = 1;
        for (const auto &t : ladder) {
            if ((size_t)t[0] * t[1] * words * 4 + slots <= (size_t)cap) {
                *tw = t[0];
                *th = t[1];
                break;
            }
        }
    }
    const size_t bits = (((size_t)(*tw) * (*th) * words * 4) + 15) & ~(size_t)15;
    *lds = (int)(bits + slots);
}

// floor(RN(a / d)) as the kernels compute it (they use the reciprocal
// sequence only where it equals IEEE division).
int host_floor_div(float a, float d)
{
    const float q = std::floor(a / d);
    if (!(q == q)) return 0;
    if (q >= 2147483647.0f) return 2147483647;
    if (q <= -2147483648.0f) return (int)0x80000000u;
    return (int)q;
}

// Extended cell grid of a plane: the union of all windows (every step from a
// pixel index to a window bound is monotone, so the first and last pixel with
// the extreme offsets bound it).  Same float operations as the kernels.
bool plane_geometry(const fg_plane_desc *d, const float4 &r, fg::PixGeom *g)
{
    const float iz = 1.0f / d->zoom, rm = d->rm, delta = d->delta;
    const float px_lo = (0.0f + 0.5f) * iz, px_hi = ((float)(d->out_w - 1) + 0.5f) * iz;
    const float py_lo = (0.0f + 0.5f) * iz, py_hi = ((float)(d->out_h - 1) + 0.5f) * iz;
    const long long I0 = host_floor_div((px_lo - r.y) - rm, delta);
    const long long I1 = host_floor_div((px_hi - r.x) + rm, delta);
    const long long J0 = host_floor_div((py_lo - r.w) - rm, delta);
    const long long J1 = host_floor_div((py_hi - r.z) + rm, delta);
    const long long W = I1 - I0 + 1, H = J1 - J0 + 1;
    if (W <= 0 || H <= 0 || W > 65535 || H > 65535) return false;
    const long long nbx = (W + fg::kBinSide - 1) / fg::kBinSide, nby = (H + fg::kBinSide - 1) / fg::kBinSide;
    if (nbx * nby >= 0x7fffffffLL) return false;
    g->Ix0 = (int)I0;
    g->Jy0 = (int)J0;
    g->W = (int)W;
    g->H = (int)H;
    g->nbx = (int)nbx;
    g->nby = (int)nby;
    return true;
}

// Bounds the splat kernel relies on, from the plane's extended cell grid and
// the offsets range r (input pixels):
//   eps     every coordinate of a candidate box (output pixels) is below
//           C_out in magnitude, so each f32 operation on it is off by at most
//           ulp(C_out)/2 <= C_out * 2^-24; the box and the fraction query take
//           fewer than ten such steps, and eps = C_out * 2^-20 + 2^-16 covers
//           them with room to spare.
//   safe_r  a grain that passes the f32 disk test of a point lies within
//           r (1 + 2^-23) + ulp(C_in) of it, and its centre lies in
//           [i delta, (i + 1) delta) up to ulp(C_in); with rm - r above
//           sixteen ulp(C_in) >= C_in * 2^-19, the points whose disk test it
//           passes have windows (f32 floor((x -+ rm) / delta), rounding
//           included) that contain its cell.  C_in bounds every coordinate
//           in input pixels.
void splat_bounds(const fg::PixelArgs &a, const fg::PixGeom &g, const float4 &r, float *eps, float *safe_r)
{
    const double maxo = std::max(std::max(std::fabs((double)r.x), std::fabs((double)r.y)),
                                 std::max(std::fabs((double)r.z), std::fabs((double)r.w)));
    const double cells = (double)std::max(std::max(std::llabs((long long)g.Ix0), std::llabs((long long)g.Ix0 + g.W)),
                                          std::max(std::llabs((long long)g.Jy0), std::llabs((long long)g.Jy0 + g.H)));
    const double pix = (double)std::max(a.out_w, a.out_h) * (double)a.inv_zoom;
    const double c_in = std::max(cells * (double)a.delta, pix) + (double)a.rm + maxo + 4.0;
    const double c_out = (c_in + maxo + (double)a.rm) * (double)a.zoom + 8.0;
    *eps = (float)(c_out * std::ldexp(1.0, -20) + std::ldexp(1.0, -16));
    const double lim = (double)a.rm - c_in * std::ldexp(1.0, -19);
    float sr = (float)lim;
    if ((double)sr > lim) sr = std::nextafter(sr, -INFINITY);
    *safe_r = sr > 0.0f ? sr : -1.0f;
}

// Heaviest-first launch order of n tiles: the caller's weights kernel fills
// TileOrder::w / ids, then sort() leaves the sorted ids in TileOrder::order.
// Workgroups start in launch order, so the heavy tiles start first and the
// light ones fill the tail.
struct TileOrder {
    uint32_t *w = nullptr, *w_sorted = nullptr;
    int *ids = nullptr, *order = nullptr;
};

int tile_order_prepare(fg_ctx *ctx, int n, TileOrder *t, hipStream_t s, const char *what)
{
    int rc = ensure_bytes(ctx, ctx->tiles, (size_t)n * 16, what);
    if (rc != FG_OK) return rc;
    size_t temp = 0;
    hipError_t e = fg::sort_tiles(nullptr, &temp, nullptr, nullptr, nullptr, nullptr, n, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if ((rc = ensure_bytes(ctx, ctx->tile_temp, temp, what)) != FG_OK) return rc;
    unsigned char *p = static_cast<unsigned char *>(ctx->tiles.p);
    t->w = reinterpret_cast<uint32_t *>(p);
    t->w_sorted = reinterpret_cast<uint32_t *>(p + (size_t)n * 4);
    t->ids = reinterpret_cast<int *>(p + (size_t)n * 8);
    t->order = reinterpret_cast<int *>(p + (size_t)n * 12);
    return FG_OK;
}

hipError_t tile_order_sort(fg_ctx *ctx, int n, const TileOrder &t, hipStream_t s)
{
    size_t temp = ctx->tile_temp.cap;
    return fg::sort_tiles(ctx->tile_temp.p, &temp, t.w, t.w_sorted, t.ids, t.order, n, s);
}

// Candidate lists of the splat's direct test for one sample table: classes
// of 1/64 output pixel up to the narrow limit rz < 0.5 and the plane's
// largest rz, 32 x 32 cells, each with its row-bucket counts.
struct ListPlan {
    int q = 32, classes = 0, ew = 0;
    // Lists for every table of at most 2048 samples (beyond, the room of
    // N * span^2 entries per class grows past 100 MB).  Short lists pay as
    // well: the splat packs the lists of consecutive grains into its 64-lane
    // steps (a 1080p RGB frame at N = 256, zoom 1, r = 0.1: 16.3 ms with
    // lists, 26.9 ms walking x-runs).
    size_t bytes(int n) const  // fill counters | (start, length) cells | row-bucket counts | entries
    {
        if (classes == 0 || n < 1 || n > 2048) return 0;
        const size_t b = 256 + (size_t)classes * q * q * (8 + 2 * fg::kRowBuckets) +
                         fg::splat_lists_entries(q, classes, ew, n) * 2;
        return (b + 255) & ~(size_t)255;
    }
};

ListPlan list_plan(const fg::PixelArgs &a, float eps)
{
    ListPlan lp;
    const float slack = 4e-3f + 2e-6f * (float)std::max(a.out_w, a.out_h);
    lp.classes = std::max(1, std::min(32, (int)std::ceil(std::min(a.rm * a.zoom + slack, 0.5f) * 64.0f)));
    lp.ew = (int)std::ceil((2.0 * (double)eps + std::ldexp(1.0, -11)) * lp.q + 0.01);
    return lp;
}

// Builds table t's lists in `at` (lp.bytes(t.n) bytes) and points t at them.
hipError_t build_lists(const ListPlan &lp, fg::SampleTable &t, unsigned char *at, int *status, hipStream_t s)
{
    uint32_t *fill = reinterpret_cast<uint32_t *>(at);
    uint2 *cs = reinterpret_cast<uint2 *>(at + 256);
    uint16_t *yx = reinterpret_cast<uint16_t *>(cs + (size_t)lp.classes * lp.q * lp.q);
    uint16_t *cl = yx + (size_t)lp.classes * lp.q * lp.q * fg::kRowBuckets;
    t.cl_cell = cs;
    t.cl_yidx = yx;
    t.cl_list = cl;
    return fg::launch_splat_lists(t, lp.q, lp.classes, lp.ew, fill, cs, yx, cl, status, s);
}

// Where a render's scratch comes from: the context's grow-only buffers
// (arena == nullptr), or a fixed arena carved front to back (the
// allocation-free entry point's static device array).
struct ScratchSource {
    unsigned char *arena = nullptr;
    size_t cap = 0, used = 0;
};

// bytes of scratch for one use: the context buffer b (grown if needed), or
// the arena's next 256-byte-aligned block (FG_ECAPACITY when it is full).
int scratch_get(fg_ctx *ctx, ScratchSource &src, DevBuf &b, size_t bytes, void **out, const char *what)
{
    if (!src.arena) {
        const int rc = ensure_bytes(ctx, b, bytes, what);
        *out = b.p;
        return rc;
    }
    const size_t off = (src.used + 255) & ~(size_t)255;
    if (off > src.cap || bytes > src.cap - off)
        return set_error(ctx, FG_ECAPACITY,
                         std::string(what) + ": the plane needs more scratch than the allocation-free entry "
                                             "point's " + std::to_string(src.cap >> 20) + " MiB");
    *out = src.arena + off;
    src.used = off + bytes;
    return FG_OK;
}

// The per-plane generation of the HBM path: Poisson thresholds, per-bin
// counts and nonempty masks, their scan, and every cell's grains written
// contiguously per bin.
// A plane's thresholds, bin counts (nbins + 1 entries) and nonempty masks,
// computed beforehand with the other planes of its frame.
struct PreCounts {
    double *T = nullptr;
    uint32_t *cnt = nullptr;
    unsigned long long *mask = nullptr;
};

struct HbmGen {
    double *T = nullptr;
    uint32_t *start = nullptr;
    unsigned long long *mask = nullptr;
    float4 *grains = nullptr;
    unsigned long long total = 0;
};

// The plane's grain total, refused when bin starts (u32 offsets) cannot
// address it: a plane of 2^32 - 256 or more grains (a large cell size or a
// bright plane) is refused before any kernel reads them (the splat's chunk
// indices run up to 63 past the last grain).
int check_total(fg_ctx *ctx, unsigned long long total, const char *what)
{
    if (total > kMaxGrains)
        return set_error(ctx, FG_ECAPACITY,
                         std::string(what) + ": the plane has " + std::to_string(total) + " grains, more than the " +
                             std::to_string(kMaxGrains) + " one render addresses; use a smaller cell size or "
                             "split the plane");
    return FG_OK;
}

int hbm_generate(fg_ctx *ctx, ScratchSource &src, fg::PixelArgs &a, const fg::PixGeom &g, int dist_kind,
                 hipStream_t s, HbmGen *out, const PreCounts *pre = nullptr)
{
    const char *what = "Pixel renderer";
    int rc;
    const long long nbins = (long long)g.nbx * g.nby;
    const size_t npx = (size_t)a.in_w * a.in_h;
    void *T_p, *cnt_p, *start_p, *mask_p, *temp_p, *grains_p;
    if (pre) {
        T_p = pre->T;
        cnt_p = pre->cnt;
        mask_p = pre->mask;
    } else {
        if ((rc = scratch_get(ctx, src, ctx->thresholds, npx * sizeof(double), &T_p, what)) != FG_OK) return rc;
        if ((rc = scratch_get(ctx, src, ctx->bin_count, (size_t)(nbins + 1) * 4, &cnt_p, what)) != FG_OK) return rc;
        if ((rc = scratch_get(ctx, src, ctx->bin_mask, (size_t)nbins * 32, &mask_p, what)) != FG_OK) return rc;
    }
    if ((rc = scratch_get(ctx, src, ctx->bin_start, (size_t)(nbins + 1) * 4, &start_p, what)) != FG_OK) return rc;
    size_t temp_bytes = 0;
    hipError_t e = fg::scan_bins(nullptr, &temp_bytes, nullptr, nullptr, (int)(nbins + 1), s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if ((rc = scratch_get(ctx, src, ctx->scan_temp, temp_bytes, &temp_p, what)) != FG_OK) return rc;
    double *T = static_cast<double *>(T_p);
    uint32_t *cnt = static_cast<uint32_t *>(cnt_p);
    uint32_t *start = static_cast<uint32_t *>(start_p);
    unsigned long long *mask = static_cast<unsigned long long *>(mask_p);
    // the plane's grain total: the bin counts summed in 64 bits (cleared by
    // a kernel: the caller may capture this into a graph)
    e = fg::launch_fill_u32(reinterpret_cast<uint32_t *>(ctx->d_sum64), 2, 0u, s);
    if (e == hipSuccess) e = fg::launch_fill_u32(cnt + nbins, 1, 0u, s);
    if (!pre) {
        if (e == hipSuccess) e = fg::launch_threshold(a, T, s);
        if (e == hipSuccess) e = fg::launch_bin_count(a, g, T, cnt, mask, s);
    }
    if (e == hipSuccess) e = fg::launch_sum_u64(cnt, nbins, ctx->d_sum64, s);
    if (e == hipSuccess) e = fg::scan_bins(temp_p, &temp_bytes, cnt, start, (int)(nbins + 1), s);
    unsigned long long total = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&total, ctx->d_sum64, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if ((rc = check_total(ctx, total, what)) != FG_OK) return rc;
    if ((rc = scratch_get(ctx, src, ctx->grains, std::max<size_t>((size_t)total, 1) * sizeof(float4), &grains_p,
                          what)) != FG_OK)
        return rc;
    float4 *grains = static_cast<float4 *>(grains_p);
    e = fg::launch_bin_grains(a, g, dist_kind, T, start, mask, grains, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    out->T = T;
    out->start = start;
    out->mask = mask;
    out->grains = grains;
    out->total = total;
    return FG_OK;
}

// Per-plane generation into HBM, then the splat-from-HBM kernel.  With an
// arena, every scratch buffer comes from it and nothing is allocated.
int run_pixel_hbm(fg_ctx *ctx, fg::PixelArgs &a, const fg::PixGeom &g, const float4 &range, int dist_kind,
                  int block, hipStream_t s, ScratchSource src = ScratchSource(), const PreCounts *pre = nullptr)
{
    const char *what = "Pixel renderer";
    int rc = FG_OK;
    // the context's buffers are shared by its calls: wait for the previous
    // user (an arena's callers are told that overlapping calls race)
    if (!src.arena && (rc = scratch_acquire(ctx, what)) != FG_OK) return rc;
    const int q = sample_buckets((uint32_t)a.n_samples);
    void *samples_p;
    if ((rc = scratch_get(ctx, src, ctx->samples, fg::sample_table_bytes(a.n_samples, q), &samples_p, what)) != FG_OK)
        return rc;
    fg::SampleTable st{};
    splat_bounds(a, g, range, &st.eps, &st.safe_r);
    mark(ctx, 0, s);
    HbmGen gen;
    if ((rc = hbm_generate(ctx, src, a, g, dist_kind, s, &gen, pre)) != FG_OK) return rc;
    uint32_t *start = gen.start;
    const float4 *grains = gen.grains;
    // heaviest-first tile order: weights and ids (in, out), sort workspace
    TileOrder order;
    void *tiles_p, *tile_temp_p;
    size_t sort_temp = 0;
    hipError_t e = fg::sort_tiles(nullptr, &sort_temp, nullptr, nullptr, nullptr, nullptr, a.n_tiles, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if ((rc = scratch_get(ctx, src, ctx->tiles, (size_t)a.n_tiles * 16, &tiles_p, what)) != FG_OK) return rc;
    if ((rc = scratch_get(ctx, src, ctx->tile_temp, sort_temp, &tile_temp_p, what)) != FG_OK) return rc;
    {
        unsigned char *p = static_cast<unsigned char *>(tiles_p);
        order.w = reinterpret_cast<uint32_t *>(p);
        order.w_sorted = reinterpret_cast<uint32_t *>(p + (size_t)a.n_tiles * 4);
        order.ids = reinterpret_cast<int *>(p + (size_t)a.n_tiles * 8);
        order.order = reinterpret_cast<int *>(p + (size_t)a.n_tiles * 12);
    }
    e = fg::launch_sample_table(a.offs, a.n_samples, a.zoom, q, samples_p, &st, s);
    // candidate lists for the direct test, one set per sample table
    const ListPlan lp = list_plan(a, st.eps);
    // (a fixed-N render touches no host container: the allocation-free
    // entry point takes this path)
    std::vector<int> &ln = ctx->level_n;
    size_t lists_total = lp.bytes(a.n_samples);
    if (a.budget) {
        ln.clear();
        for (int l = 16; l < a.n_samples; l *= 2) ln.push_back(l);
        ln.push_back(a.n_samples);
        lists_total = 0;
        for (int l : ln) lists_total += lp.bytes(l);
    }
    unsigned char *lists_at = nullptr;
    if (lists_total > 0 && lists_total < ((size_t)1 << 32)) {
        void *cells_p;
        if ((rc = scratch_get(ctx, src, ctx->cells, lists_total, &cells_p, what)) != FG_OK) return rc;
        lists_at = static_cast<unsigned char *>(cells_p);
        a.cl_q = lp.q;
        a.cl_classes = lp.classes;
        a.cl_inv_dz = (float)(2 * lp.q);
    }
    auto build_lists_for = [&](fg::SampleTable &t) -> hipError_t {
        const size_t b = lp.bytes(t.n);
        if (a.cl_q == 0 || b == 0) return hipSuccess;
        const hipError_t err = build_lists(lp, t, lists_at, a.status, s);
        lists_at += b;
        return err;
    };
    if (e == hipSuccess && !a.budget) {
        // the splat kernel reads its one table from device memory
        void *levels_p;
        if ((rc = scratch_get(ctx, src, ctx->levels, 256, &levels_p, what)) != FG_OK) return rc;
        fg::SampleTable *d_table = static_cast<fg::SampleTable *>(levels_p);
        e = build_lists_for(st);
        if (e == hipSuccess) e = fg::launch_store_table(st, d_table, s);
        a.level_tables = d_table;
    }
    if (e == hipSuccess && a.budget) {
        // adaptive sampling: one sample table per level (the first 16, 32,
        // ... samples, then all), and each tile's level
        const int nl = (int)ln.size();
        size_t bytes = 0;
        for (int l : ln) bytes += (fg::sample_table_bytes(l, sample_buckets((uint32_t)l)) + 255) & ~(size_t)255;
        const size_t head = 256 + (((size_t)nl * sizeof(fg::SampleTable) + 255) & ~(size_t)255);
        void *levels_p;
        if ((rc = scratch_get(ctx, src, ctx->levels, head + bytes + (size_t)a.n_tiles, &levels_p, what)) != FG_OK)
            return rc;
        unsigned char *base = static_cast<unsigned char *>(levels_p);
        int *d_level_n = reinterpret_cast<int *>(base);
        fg::SampleTable *d_tables = reinterpret_cast<fg::SampleTable *>(base + 256);
        unsigned char *p = base + head;
        ctx->level_tables.assign(nl, fg::SampleTable{});
        for (int i = 0; i < nl && e == hipSuccess; ++i) {
            fg::SampleTable &lt = ctx->level_tables[i];
            if (ln[i] == a.n_samples) {
                lt = st;  // the last level is the full table
            } else {
                e = fg::launch_sample_prefix(st, ln[i], sample_buckets((uint32_t)ln[i]), p, &lt, s);
                lt.eps = st.eps;  // bounds of the whole offset set hold for every prefix
                lt.safe_r = st.safe_r;
            }
            if (e == hipSuccess) e = build_lists_for(lt);
            p += (fg::sample_table_bytes(ln[i], sample_buckets((uint32_t)ln[i])) + 255) & ~(size_t)255;
        }
        uint8_t *d_tile_level = p;
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_tables, ctx->level_tables.data(), sizeof(fg::SampleTable) * nl,
                               hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(d_level_n, ln.data(), sizeof(int) * nl, hipMemcpyHostToDevice, s);
        a.level_tables = d_tables;
        a.tile_level = d_tile_level;
        if (e == hipSuccess) e = fg::launch_tile_levels(a, d_level_n, nl, d_tile_level, s);
        // the host arrays above must outlive the copies
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e == hipSuccess) e = fg::launch_pixel_tile_weights(a, g, start, order.w, order.ids, s);
    if (e == hipSuccess)
        e = fg::sort_tiles(tile_temp_p, &sort_temp, order.w, order.w_sorted, order.ids, order.order, a.n_tiles, s);
    a.tile_list = order.order;
    mark(ctx, 1, s);
    mark(ctx, 2, s);
    if (e == hipSuccess) e = fg::launch_splat_hbm(a, g, st, start, grains, block, s);
    mark(ctx, 3, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    return src.arena ? FG_OK : scratch_release(ctx, s, what);
}

// Launch arguments of one pixel-wise plane (all but the launch order and the
// tile splat kernel's offsets staging flag).
int pixel_args(fg_ctx *ctx, const fg_plane_desc *d, const float *d_lambda, const float *d_offs, float *d_out, int tw,
               int th, int lds, const float4 *d_range, int *d_status, fg::PixelArgs *out)
{
    fg::PixelArgs a{};
    a.lambda = d_lambda;
    a.offs = reinterpret_cast<const float2 *>(d_offs);
    a.out = d_out;
    a.status = d_status;
    a.cancel = ctx->d_cancel;
    a.stats = ctx->d_stats;
    a.tile_list = nullptr;
    a.off_range = d_range;
    a.seed = d->seed;
    a.log_mu = (double)d->log_mu;
    a.log_sigma = (double)d->log_sigma;
    a.in_w = (int)d->in_w;
    a.in_h = (int)d->in_h;
    a.out_w = (int)d->out_w;
    a.out_h = (int)d->out_h;
    a.n_samples = (int)d->n_samples;
    a.tile_w = tw;
    a.tile_h = th;
    a.tiles_x = (a.out_w + tw - 1) / tw;
    const long long ntiles = (long long)a.tiles_x * ((a.out_h + th - 1) / th);
    if (ntiles > 0x7fffffffLL) return set_error(ctx, FG_EINVAL, "too many tiles");
    a.n_tiles = (int)ntiles;
    a.lds_bytes = lds;
    a.words = (int)((d->n_samples + 31) / 32);
    a.inv_zoom = 1.0f / d->zoom;
    a.zoom = d->zoom;
    a.rm = d->rm;
    a.delta = d->delta;
    a.rdelta = 1.0f / d->delta;
    a.div_fast = division_fast_ok(ctx, a.delta, a.rdelta) ? 1 : 0;
    a.uscale = fg::uniform_f32_scale(0.0f, d->delta);
    float rc = d->radius_mean;  // RadiusProfile::sample for Const, then the rm clamp
    if (rc > d->rm) rc = d->rm;
    a.r_const = rc;
    a.r2_const = rc <= 0.0f ? -1.0f : rc * rc;
    a.inv_samples = 1.0f / (float)std::max<uint32_t>(d->n_samples, 1);
    a.grain_total = nullptr;
    a.grain_cap = 0;
    *out = a;
    return FG_OK;
}

// Pixel-wise render; with budget set (adaptive sampling), only the HBM splat
// formulation applies.
int do_pixelwise_budget(fg_ctx *ctx, const fg_plane_desc *d, const float *d_lambda, const float *d_offs,
                        float *d_out, hipStream_t s, bool stats, const uint16_t *budget)
{
    int variant = stats ? 1 : pixel_variant(ctx, d);
    if (budget && variant != 3)
        return set_error(ctx, FG_EINVAL, "adaptive sampling needs the HBM splat formulation (n_samples <= 2048)");
    fg::PixGeom geom{};
    float4 range{};
    hipError_t e = fg::launch_offsets_range(reinterpret_cast<const float2 *>(d_offs), (int)d->n_samples,
                                            ctx->d_range, s);
    if (e != hipSuccess) return hip_error(ctx, e, "Pixel renderer");
    if (variant == 3) {
        e = hipMemcpyAsync(&range, ctx->d_range, sizeof range, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_error(ctx, e, "Pixel renderer");
        if (!plane_geometry(d, range, &geom)) variant = 2;
    }
    int tw, th, lds;
    if (variant == 3 && !hbm_config(ctx, d, &tw, &th, &lds, 4LL * ctx->num_cus)) variant = 2;
    if (budget && variant != 3)
        return set_error(ctx, FG_EINVAL, "adaptive sampling: the plane's cell grid does not fit the HBM splat path");
    if (variant != 3) pixel_config(ctx, d, variant, &tw, &th, &lds);
    fg::PixelArgs a{};
    int rc0 = pixel_args(ctx, d, d_lambda, d_offs, d_out, tw, th, lds, ctx->d_range, ctx->d_status, &a);
    if (rc0 != FG_OK) return rc0;
    a.budget = budget;
    const size_t bits = (((size_t)tw * th * a.words * 4) + 15) & ~(size_t)15;
    // the tile splat kernel needs room for one grain record (32 B) per thread
    const bool splat = variant == 2;
    const size_t min_grains = 256;
    a.offs_in_lds = (splat && d->n_samples <= 2048 &&
                     bits + (size_t)d->n_samples * 8 + min_grains * 32 <= (size_t)lds) ? 1 : 0;
    if (splat && bits + (a.offs_in_lds ? (size_t)d->n_samples * 8 : 0) + min_grains * 32 > (size_t)lds)
        return set_error(ctx, FG_EINVAL, "sample count too large for one pixel-wise tile");
    if (variant == 3) return run_pixel_hbm(ctx, a, geom, range, (int)d->dist_kind, splat_hbm_block(ctx, d), s);
    mark(ctx, 0, s);
    e = fg::launch_lambda_check(d_lambda, (long long)d->in_w * d->in_h, a.status, s);
    if (e != hipSuccess) return hip_error(ctx, e, "Pixel renderer");
    mark(ctx, 1, s);
    mark(ctx, 2, s);
    e = fg::launch_pixelwise(a, (int)d->dist_kind, variant == 1 ? 1 : 0, stats, s);
    mark(ctx, 3, s);
    if (e != hipSuccess) return hip_error(ctx, e, "Pixel renderer");
    return FG_OK;
}

int do_pixelwise(fg_ctx *ctx, const fg_plane_desc *d, const float *d_lambda, const float *d_offs,
                 float *d_out, hipStream_t s, bool stats)
{
    return do_pixelwise_budget(ctx, d, d_lambda, d_offs, d_out, s, stats, nullptr);
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The planes of one frame (one description, n λ planes): on the HBM path one
// counting pass serves all of them (bin_count_planes_kernel), then each plane
// draws its grains and is splatted as a single plane is.  Other formulations
// render the planes one by one.
int do_pixelwise_planes(fg_ctx *ctx, const fg_plane_desc *d, int n, const float *const *lams, const float *d_offs,
                        float *const *outs, hipStream_t s)
{
    const char *what = "Pixel renderer";
    int variant = pixel_variant(ctx, d);
    float4 range{};
    fg::PixGeom geom{};
    int tw = 0, th = 0, lds = 0;
    bool joint = n > 1 && n <= fg::kFramePlanes && variant == 3;
    if (joint) {
        hipError_t e = fg::launch_offsets_range(reinterpret_cast<const float2 *>(d_offs), (int)d->n_samples,
                                                ctx->d_range, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&range, ctx->d_range, sizeof range, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_error(ctx, e, what);
        joint = plane_geometry(d, range, &geom) && hbm_config(ctx, d, &tw, &th, &lds, 4LL * ctx->num_cus);
    }
    if (!joint) {
        for (int i = 0; i < n; ++i) {
            const int rc = do_pixelwise(ctx, d, lams[i], d_offs, outs[i], s, false);
            if (rc != FG_OK) return rc;
        }
        return FG_OK;
    }
    fg::PixelArgs a[fg::kFramePlanes];
    for (int i = 0; i < n; ++i) {
        const int rc = pixel_args(ctx, d, lams[i], d_offs, outs[i], tw, th, lds, ctx->d_range, ctx->d_status, &a[i]);
        if (rc != FG_OK) return rc;
    }
    // The frame's pipeline: thresholds and the joint count, every plane's
    // scan, one host read of all the grain totals (their sizes), then each
    // plane's draws and tile order, one sample table and its candidate lists
    // for all the planes (they share the offsets), and the splats.  Every
    // plane has its own bin starts, grains and tile order, so nothing waits
    // for a plane's splat; the previous call's kernels may still read the
    // context's buffers, hence the one acquire.
    int rc = scratch_acquire(ctx, what);
    if (rc != FG_OK) return rc;
    const int kind = (int)d->dist_kind, block = splat_hbm_block(ctx, d);
    const long long nbins = (long long)geom.nbx * geom.nby;
    const size_t t_b = align256((size_t)d->in_w * d->in_h * sizeof(double));
    const size_t c_b = align256((size_t)(nbins + 1) * 4), m_b = align256((size_t)nbins * 32);
    const size_t per_plane = t_b + 2 * c_b + m_b;  // thresholds | counts | starts | masks
    if ((rc = ensure_bytes(ctx, ctx->plane_counts, 256 + (size_t)n * per_plane, what)) != FG_OK) return rc;
    unsigned char *base = static_cast<unsigned char *>(ctx->plane_counts.p);
    unsigned long long *d_tot = reinterpret_cast<unsigned long long *>(base);  // the planes' grain totals
    uint32_t *start[fg::kFramePlanes];
    PreCounts pre[fg::kFramePlanes];
    fg::FramePlanes fp{};
    for (int i = 0; i < n; ++i) {
        unsigned char *b = base + 256 + (size_t)i * per_plane;
        pre[i].T = reinterpret_cast<double *>(b);
        pre[i].cnt = reinterpret_cast<uint32_t *>(b + t_b);
        start[i] = reinterpret_cast<uint32_t *>(b + t_b + c_b);
        pre[i].mask = reinterpret_cast<unsigned long long *>(b + t_b + 2 * c_b);
        fp.T[i] = pre[i].T;
        fp.cnt[i] = pre[i].cnt;
        fp.mask[i] = pre[i].mask;
    }
    // the rest of the scratch, sized before any kernel of this call runs
    // (growing a buffer frees the old one)
    size_t scan_bytes = 0;
    hipError_t e = fg::scan_bins(nullptr, &scan_bytes, nullptr, nullptr, (int)(nbins + 1), s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    const int q = sample_buckets((uint32_t)d->n_samples);
    fg::SampleTable st{};
    splat_bounds(a[0], geom, range, &st.eps, &st.safe_r);
    const ListPlan lp = list_plan(a[0], st.eps);
    const size_t lists_total = lp.bytes((int)d->n_samples);
    const bool lists = lists_total > 0 && lists_total < ((size_t)1 << 32);
    // the persistent splat's plane records, (plane, tile) items, their
    // weights and the item counter
    const int n_items = n * a[0].n_tiles;
    const size_t rec_b = align256(sizeof(fg::BatchPlane) * fg::kFramePlanes);
    const size_t items_b = align256((size_t)n_items * sizeof(int2)), w_b = align256((size_t)n_items * 4);
    if ((rc = ensure_bytes(ctx, ctx->scan_temp, scan_bytes, what)) != FG_OK) return rc;
    if ((rc = ensure_bytes(ctx, ctx->tiles, rec_b + items_b + w_b + 256, what)) != FG_OK) return rc;
    if ((rc = ensure_bytes(ctx, ctx->samples, fg::sample_table_bytes((int)d->n_samples, q), what)) != FG_OK) return rc;
    if ((rc = ensure_bytes(ctx, ctx->levels, 256, what)) != FG_OK) return rc;
    if (lists && (rc = ensure_bytes(ctx, ctx->cells, lists_total, what)) != FG_OK) return rc;
    mark(ctx, 0, s);
    for (int i = 0; i < n && e == hipSuccess; ++i) e = fg::launch_threshold(a[i], pre[i].T, s);
    if (e == hipSuccess) e = fg::launch_bin_count_planes(a[0], geom, fp, n, s);
    // totals (cleared by a kernel, as the single-plane path does) and scans
    if (e == hipSuccess) e = fg::launch_fill_u32(reinterpret_cast<uint32_t *>(d_tot), 2 * n, 0u, s);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        e = fg::launch_fill_u32(pre[i].cnt + nbins, 1, 0u, s);
        if (e == hipSuccess) e = fg::launch_sum_u64(pre[i].cnt, nbins, d_tot + i, s);
        if (e == hipSuccess)
            e = fg::scan_bins(ctx->scan_temp.p, &scan_bytes, pre[i].cnt, start[i], (int)(nbins + 1), s);
    }
    unsigned long long total[fg::kFramePlanes] = {};
    if (e == hipSuccess) e = hipMemcpyAsync(total, d_tot, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    // the grain array of all the planes (nothing of this call runs now)
    size_t goff[fg::kFramePlanes], gsum = 0;
    for (int i = 0; i < n; ++i) {
        if ((rc = check_total(ctx, total[i], what)) != FG_OK) return rc;
        goff[i] = gsum;
        gsum += (size_t)total[i];
    }
    if ((rc = ensure_bytes(ctx, ctx->grains, std::max<size_t>(gsum, 1) * sizeof(float4), what)) != FG_OK) return rc;
    float4 *grains = static_cast<float4 *>(ctx->grains.p);
    // one sample table and its candidate lists, which the splat reads from
    // device memory
    e = fg::launch_sample_table(a[0].offs, a[0].n_samples, a[0].zoom, q, ctx->samples.p, &st, s);
    if (e == hipSuccess && lists) e = build_lists(lp, st, static_cast<unsigned char *>(ctx->cells.p), a[0].status, s);
    fg::SampleTable *d_table = static_cast<fg::SampleTable *>(ctx->levels.p);
    if (e == hipSuccess) e = fg::launch_store_table(st, d_table, s);
    fg::FramePlaneRecords rec{};
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        e = fg::launch_bin_grains(a[i], geom, kind, pre[i].T, start[i], pre[i].mask, grains + goff[i], s);
        if (lists) {
            a[i].cl_q = lp.q;
            a[i].cl_classes = lp.classes;
            a[i].cl_inv_dz = (float)(2 * lp.q);
        }
        fg::BatchPlane &bp = rec.p[i];
        bp.a = a[i];
        bp.geo = geom;
        bp.st = st;
        bp.bin_start = start[i];
        bp.grains = grains + goff[i];
        bp.item_base = (long long)i * a[i].n_tiles;
    }
    // the three splats as one persistent launch over the frame's (plane,
    // tile) items, heaviest first
    unsigned char *tb = static_cast<unsigned char *>(ctx->tiles.p);
    fg::BatchPlane *d_planes = reinterpret_cast<fg::BatchPlane *>(tb);
    int2 *items = reinterpret_cast<int2 *>(tb + rec_b);
    uint32_t *w = reinterpret_cast<uint32_t *>(tb + rec_b + items_b);
    uint32_t *counter = reinterpret_cast<uint32_t *>(tb + rec_b + items_b + w_b);
    if (e == hipSuccess) e = fg::launch_store_frame_planes(rec, n, d_planes, s);
    if (e == hipSuccess) e = fg::launch_batch_weights(d_planes, n, w, n_items, s);
    if (e == hipSuccess) e = fg::launch_batch_order(d_planes, n, w, n_items, items, nullptr, s);
    if (e == hipSuccess) e = fg::launch_fill_u32(counter, 1, 0u, s);
    const int grid =
        std::max(1, std::min(n_items, ctx->num_cus * fg::splat_batch_blocks_per_cu(block, a[0].lds_bytes)));
    mark(ctx, 1, s);
    mark(ctx, 2, s);
    if (e == hipSuccess) e = fg::launch_splat_batch(d_planes, items, n_items, counter, grid, block, a[0].lds_bytes, s);
    mark(ctx, 3, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    return scratch_release(ctx, s, what);
}

// Grain-wise formulation: 1 the reference's loop order over all (grain,
// sample) pairs, 2 runs of the sorted sample order (automatic; needs N <= 2048).
int grain_variant(fg_ctx *ctx, const fg_plane_desc *d)
{
    if (launch_cfg(ctx, d, 1).grain_variant == 1 || d->n_samples > kHbmMaxSamples) return 1;
    return 2;
}

int grain_block(const fg_ctx *ctx, const fg_plane_desc *d)
{
    return launch_cfg(ctx, d, 1).grain_block == 256 ? 256 : 512;
}

// Sorted grain-wise kernel: the configured tile, else the largest tile of the
// ladder whose bitset takes at most 32 KB; queue entries address at most
// 128 x 256 tiles.  Returns false when no tile fits the LDS limit.
bool sorted_grain_config(fg_ctx *ctx, const fg_plane_desc *d, int *tw, int *th, int *lds)
{
    static const int ladder[][2] = {{128, 64}, {64, 64}, {64, 32}, {32, 32}, {32, 16}, {16, 16},
                                    {16, 8},   {8, 8},   {8, 4},   {4, 4},   {2, 2},   {1, 1}};
    const fg_launch_config &cfg = launch_cfg(ctx, d, 1);
    const int q = sample_buckets(d->n_samples);
    const size_t words = (d->n_samples + 31) / 32;
    const size_t limit = (size_t)ctx->max_lds - fg::grainwise_sorted_static_lds(grain_block(ctx, d));
    auto fits = [&](int w, int h) { return fg::grainwise_sorted_lds(w, h, (int)d->n_samples, q) <= limit; };
    if (cfg.grain_tile_w > 0 && cfg.grain_tile_h > 0 && cfg.grain_tile_w <= 128 &&
        cfg.grain_tile_h <= 256 && fits(cfg.grain_tile_w, cfg.grain_tile_h)) {
        *tw = cfg.grain_tile_w;
        *th = cfg.grain_tile_h;
    } else {
        size_t i = 0, n = sizeof ladder / sizeof ladder[0];
        while (i < n && !((size_t)ladder[i][0] * ladder[i][1] * words * 4 <= 32768 && fits(ladder[i][0], ladder[i][1])))
            ++i;
        if (i == n) return false;
        // smaller tiles until the plane has at least four per CU, so the
        // heaviest tile does not set the plane's time (512x512, N = 16: 32
        // tiles of 128x64 take 8 ms)
        auto tiles = [&](size_t k) {
            return (long long)((d->out_w + ladder[k][0] - 1) / ladder[k][0]) *
                   (long long)((d->out_h + ladder[k][1] - 1) / ladder[k][1]);
        };
        while (i + 1 < n && tiles(i) < 4LL * ctx->num_cus && ladder[i + 1][0] * ladder[i + 1][1] >= 64 &&
               fits(ladder[i + 1][0], ladder[i + 1][1]))
            ++i;
        *tw = ladder[i][0];
        *th = ladder[i][1];
    }
    *lds = (int)fg::grainwise_sorted_lds(*tw, *th, (int)d->n_samples, q);
    return true;
}

int do_grainwise(fg_ctx *ctx, const fg_plane_desc *d, const float *d_lambda, const float *d_offs,
                 float *d_out, hipStream_t s, bool stats)
{
    int variant = stats ? 1 : grain_variant(ctx, d);
    int tw, th, lds, gpt = 0;
    if (variant == 2 && !sorted_grain_config(ctx, d, &tw, &th, &lds)) variant = 1;
    if (variant == 1) {
        grain_config(ctx, d, &tw, &th, &lds, &gpt);
        if (lds > ctx->max_lds)
            return set_error(ctx, FG_EINVAL, "sample count too large for one grain-wise tile");
    }
    fg::GrainArgs a{};
    a.lambda = d_lambda;
    a.offs = reinterpret_cast<const float2 *>(d_offs);
    a.out = d_out;
    a.status = ctx->d_status;
    a.cancel = ctx->d_cancel;
    a.stats = ctx->d_stats;
    a.seed = d->seed;
    a.log_mu = (double)d->log_mu;
    a.log_sigma = (double)d->log_sigma;
    a.in_w = (int)d->in_w;
    a.in_h = (int)d->in_h;
    a.out_w = (int)d->out_w;
    a.out_h = (int)d->out_h;
    a.n_samples = (int)d->n_samples;
    a.words = (int)((d->n_samples + 31) / 32);
    a.tile_w = tw;
    a.tile_h = th;
    a.tiles_x = (a.out_w + tw - 1) / tw;
    const long long ntiles = (long long)a.tiles_x * ((a.out_h + th - 1) / th);
    if (ntiles > 0x7fffffffLL) return set_error(ctx, FG_EINVAL, "too many tiles");
    a.n_tiles = (int)ntiles;
    a.grains_per_thread = gpt;
    a.lds_bytes = lds;
    a.zoom = d->zoom;
    a.rm = d->rm;
    a.r_const = d->radius_mean;
    a.inv_samples = 1.0f / (float)std::max<uint32_t>(d->n_samples, 1);
    a.r_out_max = d->rm * d->zoom;
    a.off_range = ctx->d_range;
    const char *what = "Grain renderer";
    if (variant == 2) {
        // the sample order lives in the context's scratch
        int rc = scratch_acquire(ctx, what);
        if (rc != FG_OK) return rc;
        const int q = sample_buckets(d->n_samples);
        if ((rc = ensure_bytes(ctx, ctx->samples, fg::sample_table_bytes(a.n_samples, q), what)) != FG_OK)
            return rc;
        const size_t npx = (size_t)a.in_w * a.in_h;
        if ((rc = ensure_bytes(ctx, ctx->pix_state, npx * 36 + 32, what)) != FG_OK) return rc;
        unsigned char *ps = static_cast<unsigned char *>(ctx->pix_state.p);
        a.pix_s01 = reinterpret_cast<ulonglong2 *>(ps);
        a.pix_s23 = reinterpret_cast<ulonglong2 *>(ps + npx * 16);
        a.pix_q = reinterpret_cast<uint32_t *>(ps + npx * 32);
        TileOrder order;
        if ((rc = tile_order_prepare(ctx, a.n_tiles, &order, s, what)) != FG_OK) return rc;
        fg::SampleTable st{};
        mark(ctx, 0, s);
        hipError_t e = fg::launch_sample_table(a.offs, a.n_samples, 1.0f, q, ctx->samples.p, &st, s);
        // constant radius with a narrow window (2R + 1/Q <= 0.35 output px):
        // candidate lists per cell of the grain's fractional target, about
        // N (2R + 1/Q)^2 samples each instead of a run of N (2R + 1/Q)
        const int cq = 64;
        const float r_out = std::min(d->radius_mean, d->rm) * d->zoom;
        if (e == hipSuccess && d->dist_kind == 0 && r_out > 0.0f && 2.0f * r_out + 1.0f / (float)cq <= 0.35f) {
            const size_t nstart = (size_t)cq * cq + 1;
            if ((rc = ensure_bytes(ctx, ctx->cells, nstart * 4 + (size_t)cq * cq * a.n_samples * 2 + 64, what)) !=
                FG_OK)
                return rc;
            uint32_t *cs = static_cast<uint32_t *>(ctx->cells.p);
            uint16_t *cl = reinterpret_cast<uint16_t *>(cs + nstart);
            const float c_base = (float)std::max(a.out_w, a.out_h) + a.r_out_max + 8.0f;
            e = fg::launch_cell_lists(st, a.n_samples, cq, r_out, c_base, cs, cl, s);
            a.cell_start = cs;
            a.cell_list = cl;
            a.cell_q = cq;
        }
        if (e == hipSuccess) e = fg::launch_pixel_state(a, s);
        if (e == hipSuccess) e = fg::launch_grain_tile_weights(a, st, order.w, order.ids, s);
        if (e == hipSuccess) e = tile_order_sort(ctx, a.n_tiles, order, s);
        a.tile_list = order.order;
        mark(ctx, 1, s);
        mark(ctx, 2, s);
        if (e == hipSuccess) e = fg::launch_grainwise_sorted(a, st, (int)d->dist_kind, grain_block(ctx, d), s);
        mark(ctx, 3, s);
        if (e != hipSuccess) return hip_error(ctx, e, what);
        return scratch_release(ctx, s, what);
    }
    hipError_t e = fg::launch_offsets_range(a.offs, a.n_samples, ctx->d_range, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    mark(ctx, 0, s);
    if (stats) e = fg::launch_grain_count(a, (int)d->dist_kind, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    mark(ctx, 1, s);
    mark(ctx, 2, s);
    e = fg::launch_grainwise(a, (int)d->dist_kind, stats, s);
    mark(ctx, 3, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    return FG_OK;
}

// Device status bits to a status code (no bits: FG_OK).
int status_error(fg_ctx *ctx, int status, const char *what)
{
    if (status & fg::kStatusCancelled) return set_error(ctx, FG_ECANCELLED, "cancelled");
    if (status & fg::kStatusNeighbourhood)
        return set_error(ctx, FG_ENEIGHBOURHOOD,
                         std::string(what) +
                             ": the cell window of one output pixel does not fit in LDS; "
                             "use a larger cell size");
    if (status & fg::kStatusRegion)
        return set_error(ctx, FG_EINVAL,
                         std::string(what) + ": the cell region of one tile exceeds 65535 cells per side; "
                                             "use a larger cell size");
    if (status & fg::kStatusLambda)
        return set_error(ctx, FG_EINVAL,
                         std::string(what) + ": lambda values must be finite and at most 1e6 (MAX_LAMBDA); "
                                             "NaN, +inf and larger values were skipped");
    if (status & fg::kStatusOverflow)
        return set_error(ctx, FG_ECAPACITY,
                         std::string(what) + ": a plane's grain total reached 2^32; use a smaller cell size");
    if (status & fg::kStatusCapacity)
        return set_error(ctx, FG_ECAPACITY,
                         std::string(what) + ": planned capacity exceeded (a batch's grains, or the splat's "
                                             "candidate lists); for a batch, create it again");
    if (status != 0) return set_error(ctx, FG_EHIP, std::string(what) + ": device status " + std::to_string(status));
    return FG_OK;
}

void set_cancel(fg_ctx *ctx, int on) { __atomic_store_n(ctx->h_cancel, on, __ATOMIC_SEQ_CST); }

// Waits for stream s; while it runs, polls the caller's cancel callback and
// raises the context's cancel flag when it asks, so that the integrators'
// remaining tiles return at once (src/pixelwise.rs:28 checks per row).
hipError_t wait_polling(fg_ctx *ctx, hipStream_t s, fg_cancel_fn cancel, void *user)
{
    if (!cancel) return hipStreamSynchronize(s);
    bool raised = false;
    for (;;) {
        const hipError_t q = hipStreamQuery(s);
        if (q != hipErrorNotReady) return q;
        if (!raised && cancel(user)) {
            set_cancel(ctx, 1);
            raised = true;
        }
        struct timespec ts = {0, 200000};  // 0.2 ms
        nanosleep(&ts, nullptr);
    }
}

int read_status(fg_ctx *ctx, hipStream_t s, const char *what)
{
    int status = 0;
    hipError_t e = hipMemcpyAsync(&status, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if (status != 0) {
        e = hipMemsetAsync(ctx->d_status, 0, sizeof(int), s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_error(ctx, e, what);
    }
    return status_error(ctx, status, what);
}

bool finite_pairs(const float *xy, uint32_t n)
{
    for (uint32_t k = 0; k < 2 * n; ++k)
        if (!std::isfinite(xy[k])) return false;
    return true;
}

using DeviceFn = int (*)(fg_ctx *, const fg_plane_desc *, const float *, const float *, float *,
                         hipStream_t, bool);

int render_host(fg_ctx *ctx, const fg_plane_desc *d, const float *lambda, const float *offs,
                float *out, fg_cancel_fn cancel, void *user, DeviceFn fn, const char *what)
{
    int rc = validate(ctx, d);
    if (rc != FG_OK) return rc;
    if (!lambda || !offs || !out) return set_error(ctx, FG_EINVAL, "null buffer");
    if (!finite_pairs(offs, d->n_samples))
        return set_error(ctx, FG_EINVAL, "offsets must be finite");
    if (cancel && cancel(user)) return set_error(ctx, FG_ECANCELLED, "cancelled");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    const size_t nin = (size_t)d->in_w * d->in_h, nout = (size_t)d->out_w * d->out_h;
    if ((rc = ensure(ctx, &ctx->d_lambda, &ctx->cap_lambda, nin, what)) != FG_OK) return rc;
    if ((rc = ensure(ctx, &ctx->d_offs, &ctx->cap_offs, 2 * (size_t)d->n_samples, what)) != FG_OK)
        return rc;
    if ((rc = ensure(ctx, &ctx->d_out, &ctx->cap_out, nout, what)) != FG_OK) return rc;
    hipStream_t s = ctx->stream;
    set_cancel(ctx, 0);
    e = hipMemcpyAsync(ctx->d_lambda, lambda, nin * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->d_offs, offs, (size_t)d->n_samples * 8, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    rc = fn(ctx, d, ctx->d_lambda, ctx->d_offs, ctx->d_out, s, false);
    if (rc != FG_OK) return rc;
    e = wait_polling(ctx, s, cancel, user);
    if (e == hipSuccess) e = hipMemcpyAsync(out, ctx->d_out, nout * 4, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    rc = read_status(ctx, s, what);
    set_cancel(ctx, 0);
    return rc;
}

int stats_device(fg_ctx *ctx, const fg_plane_desc *d, const float *d_lambda, const float *d_offs,
                 float *d_out, uint64_t *counters, int n_counters, void *stream, DeviceFn fn, const char *what)
{
    int rc = validate(ctx, d);
    if (rc != FG_OK) return rc;
    if (!d_lambda || !d_offs || !d_out || !counters) return set_error(ctx, FG_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->d_stats, 0, 4 * sizeof(unsigned long long), s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    rc = fn(ctx, d, d_lambda, d_offs, d_out, s, true);
    if (rc != FG_OK) return rc;
    unsigned long long h[4] = {0, 0, 0, 0};
    e = hipMemcpyAsync(h, ctx->d_stats, sizeof h, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    rc = read_status(ctx, s, what);
    for (int i = 0; i < n_counters; ++i) counters[i] = h[i];
    return rc;
}

int validate_frame(fg_ctx *ctx, const fg_frame_desc *fd)
{
    if (!ctx) return FG_EINVAL;
    if (!fd) return set_error(ctx, FG_EINVAL, "null frame description");
    const int rc = validate(ctx, &fd->plane);
    if (rc != FG_OK) return rc;
    if (fd->color_mode > 1) return set_error(ctx, FG_EINVAL, "unknown colour mode");
    if (fd->algo > 1) return set_error(ctx, FG_EINVAL, "unknown algorithm");
    if (!std::isfinite(fd->inv_e_pi_r2) || !(fd->inv_e_pi_r2 >= 0.0f))
        return set_error(ctx, FG_EINVAL, "inv_e_pi_r2 must be finite and >= 0");
    return FG_OK;
}


// One frame on the device (src/lib.rs:141-165 with src/color.rs:47-114):
// split, then per rendered plane the lambda pre-pass and the integrator, then
// the merge.  The frame scratch holds the three split planes, one lambda
// plane (reused: the planes run in stream order), the max word and, without
// d_planes, the rendered planes.
int frame_device(fg_ctx *ctx, const fg_frame_desc *fd, const float *d_rgb, const float *d_offs, uint8_t *d_rgb8,
                 float *d_planes, hipStream_t s, fg_cancel_fn cancel, void *user)
{
    const char *what = "Frame renderer";
    const fg_plane_desc &d = fd->plane;
    const size_t nin = (size_t)d.in_w * d.in_h, nout = (size_t)d.out_w * d.out_h;
    const int nr = fd->color_mode == FG_COLOR_RGB ? 3 : 1;
    const size_t in_b = align256(nin * 4), out_b = align256(nout * 4);
    // RGB pixel-wise frames keep all three lambda planes for one joint render
    const bool joint = nr == 3 && fd->algo == FG_ALGO_PIXEL;
    const size_t bytes = (joint ? 6 : 4) * in_b + 256 + (d_planes ? 0 : (size_t)nr * out_b);
    // an earlier call's kernels may still read the scratch (on another stream)
    int rc = scratch_acquire(ctx, what);
    if (rc != FG_OK) return rc;
    if ((rc = ensure_bytes(ctx, ctx->frame, bytes, what)) != FG_OK) return rc;
    unsigned char *base = static_cast<unsigned char *>(ctx->frame.p);
    float *sp[3] = {reinterpret_cast<float *>(base), reinterpret_cast<float *>(base + in_b),
                    reinterpret_cast<float *>(base + 2 * in_b)};
    const size_t lam_at = joint ? 6 * in_b : 4 * in_b;  // after the split planes and the lambda planes
    float *lam3[3] = {reinterpret_cast<float *>(base + 3 * in_b), reinterpret_cast<float *>(base + 4 * in_b),
                      reinterpret_cast<float *>(base + 5 * in_b)};
    float *lam = lam3[0];
    unsigned *word = reinterpret_cast<unsigned *>(base + lam_at);
    float *outp = d_planes ? d_planes : reinterpret_cast<float *>(base + lam_at + 256);
    hipError_t e = fg::launch_split((int)fd->color_mode, d_rgb, (long long)nin, sp[0], sp[1], sp[2], s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if (joint) {
        if (cancel && cancel(user)) return set_error(ctx, FG_ECANCELLED, "cancelled");
        for (int i = 0; i < 3 && e == hipSuccess; ++i)
            e = fg::launch_lambda(sp[i], (long long)nin, fd->inv_e_pi_r2, word, nullptr, lam3[i], s);
        if (e != hipSuccess) return hip_error(ctx, e, what);
        float *outs[3] = {outp, outp + nout, outp + 2 * nout};
        const float *lams[3] = {lam3[0], lam3[1], lam3[2]};
        if ((rc = do_pixelwise_planes(ctx, &d, 3, lams, d_offs, outs, s)) != FG_OK) return rc;
    } else {
        for (int i = 0; i < nr; ++i) {
            if (cancel && cancel(user)) return set_error(ctx, FG_ECANCELLED, "cancelled");
            e = fg::launch_lambda(sp[i], (long long)nin, fd->inv_e_pi_r2, word, nullptr, lam, s);
            if (e != hipSuccess) return hip_error(ctx, e, what);
            rc = fd->algo == FG_ALGO_PIXEL ? do_pixelwise(ctx, &d, lam, d_offs, outp + (size_t)i * nout, s, false)
                                           : do_grainwise(ctx, &d, lam, d_offs, outp + (size_t)i * nout, s, false);
            if (rc != FG_OK) return rc;
        }
    }
    if (fd->color_mode == FG_COLOR_RGB)
        e = fg::launch_merge(FG_COLOR_RGB, outp, outp + nout, outp + 2 * nout, (int)d.out_w, (int)d.out_h,
                             (int)d.out_w, (int)d.out_h, d_rgb8, s);
    else
        e = fg::launch_merge(FG_COLOR_LUMA, outp, sp[1], sp[2], (int)d.out_w, (int)d.out_h, (int)d.in_w,
                             (int)d.in_h, d_rgb8, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    return scratch_release(ctx, s, what);
}

}  // namespace

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
struct fg_batch {
    fg_ctx *ctx = nullptr;
    int n = 0;
    std::vector<void *> allocs;
    std::vector<fg::BatchPlane> planes;
    std::vector<ListPlan> list_plans;       // per plane: the splat's candidate lists
    std::vector<unsigned char *> list_mem;  // per plane: their buffer, or nullptr
    std::vector<int> kind;             // radius law per plane
    std::vector<double *> T;           // thresholds per plane
    std::vector<long long> bin_off;    // first bin of each plane in the batch's bins
    std::vector<void *> samples;       // sample table buffer per plane
    std::vector<int> qb;               // sample buckets per plane
    // runs of consecutive planes with one description and offset range (the
    // planes of one image): counted in one pass, each cell hashed once
    std::vector<std::pair<int, int>> count_groups;  // (first plane, planes)
    // the generation passes, each one launch over all planes (its offset
    // and pointer arrays in gen_mem)
    fg::BatchGen gen{};
    unsigned char *gen_mem = nullptr;
    fg::BatchPlane *d_planes = nullptr;
    float4 *ranges = nullptr;
    int *status = nullptr;
    uint32_t *cnt = nullptr, *start = nullptr, *partials = nullptr, *weights = nullptr, *counter = nullptr;
    unsigned long long *total64 = nullptr;  // the batch's grain total in 64 bits
    unsigned long long *mask = nullptr;
    int2 *items = nullptr;
    float4 *grains = nullptr;
    long long total_bins = 0;
    int n_items = 0, grid = 0, block = 512, lds = 0;
    uint32_t cap = 0;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int batch_alloc(fg_batch *b, void **p, size_t bytes, const char *what)
{
    const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return hip_error(b->ctx, e, what);
    }
    b->allocs.push_back(*p);
    return FG_OK;
}

void batch_free(fg_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    for (void *p : b->allocs) (void)hipFree(p);
    b->allocs.clear();
    b->exec = nullptr;
    b->graph = nullptr;
    b->capture_stream = nullptr;
}

// Generation of every plane, then the batch scan (the counting half).
hipError_t batch_enqueue_counts(fg_batch *b, hipStream_t s)
{
    hipError_t e = fg::launch_fill_u32(reinterpret_cast<uint32_t *>(b->status), 1, 0u, s);
    if (e == hipSuccess) e = fg::launch_fill_u32(reinterpret_cast<uint32_t *>(b->total64), 2, 0u, s);
    if (e == hipSuccess) e = fg::launch_batch_threshold(b->gen, s);
    // counts per group of planes: a launch per group keeps the plane's
    // arguments in the kernel's argument segment (one launch over all
    // groups, each wave searching for its group, was 10% slower)
    for (const auto &grp : b->count_groups) {
        if (e != hipSuccess) break;
        const int p0 = grp.first, np = grp.second;
        fg::FramePlanes fp{};
        for (int k = 0; k < np; ++k) {
            const int p = p0 + k;
            fp.T[k] = b->T[p];
            fp.cnt[k] = b->cnt + b->bin_off[p];
            fp.mask[k] = b->mask + 4 * b->bin_off[p];
        }
        const fg::BatchPlane &bp = b->planes[p0];
        e = fg::launch_bin_count_planes(bp.a, bp.geo, fp, np, s);
    }
    if (e == hipSuccess) e = fg::launch_sum_u64(b->cnt, b->total_bins, b->total64, s);
    if (e == hipSuccess) e = fg::launch_scan_u32(b->cnt, b->start, b->total_bins, b->partials, s);
    return e;
}

// The whole pipeline of one run.
hipError_t batch_enqueue(fg_batch *b, hipStream_t s)
{
    hipError_t e = batch_enqueue_counts(b, s);
    if (e == hipSuccess) e = fg::launch_capacity_check(b->total64, b->cap, b->status, s);
    if (e == hipSuccess) e = fg::launch_batch_bin_grains(b->gen, s);
    if (e == hipSuccess) e = fg::launch_batch_sample_tables(b->gen, s);
    if (e == hipSuccess) e = fg::launch_batch_weights(b->d_planes, b->n, b->weights, b->n_items, s);
    if (e == hipSuccess) e = fg::launch_batch_order(b->d_planes, b->n, b->weights, b->n_items, b->items, nullptr, s);
    if (e == hipSuccess) e = fg::launch_fill_u32(b->counter, 1, 0u, s);
    if (e == hipSuccess)
        e = fg::launch_splat_batch(b->d_planes, b->items, b->n_items, b->counter, b->grid, b->block, b->lds, s);
    return e;
}

// The offset and pointer arrays of the batched generation passes, laid out
// in one device buffer (b->gen_mem) and described by b->gen.
int batch_gen_setup(fg_batch *b, const char *what, hipStream_t s)
{
    const int n = b->n;
    std::vector<unsigned char> host;
    std::vector<std::pair<size_t, const void **>> fixups;  // (offset in gen_mem, field of b->gen)
    auto put = [&](const void *src, size_t bytes, const void **field) {
        const size_t off = (host.size() + 15) & ~(size_t)15;
        host.resize(off + std::max<size_t>(bytes, 16));
        if (bytes) std::memcpy(host.data() + off, src, bytes);
        fixups.push_back({off, field});
    };
    // the sample tables' launch has a workgroup row per plane
    if (n > 65535) return set_error(b->ctx, FG_EINVAL, "a batch holds at most 65535 planes");
    fg::BatchGen &g = b->gen;
    g = fg::BatchGen{};
    g.n = n;
    // thresholds: kBlock pixels a workgroup
    std::vector<long long> thr(n + 1, 0), bins(n + 1, 0), lists(n + 1, 0);
    std::vector<double *> T(n);
    std::vector<int4> plan(n);
    std::vector<uint32_t *> fill(n);
    int max_samples = 1;
    for (int p = 0; p < n; ++p) {
        const fg::BatchPlane &bp = b->planes[p];
        const long long px = (long long)bp.a.in_w * bp.a.in_h;
        thr[p + 1] = thr[p] + (px + fg::kBlock - 1) / fg::kBlock;
        bins[p + 1] = bins[p] + (long long)bp.geo.nbx * bp.geo.nby;
        T[p] = b->T[p];
        max_samples = std::max(max_samples, bp.a.n_samples);
        const ListPlan &lp = b->list_plans[p];
        long long lb = 0;
        plan[p] = make_int4(0, 0, 0, 0);
        fill[p] = nullptr;
        if (b->list_mem[p]) {
            if (lp.q < 1 || lp.q * lp.q > 65536 || lp.classes < 1) return set_error(b->ctx, FG_EINVAL, "batch list plan");
            plan[p] = make_int4(lp.q, lp.classes, lp.ew, 0);
            fill[p] = reinterpret_cast<uint32_t *>(b->list_mem[p]);
            lb = (long long)lp.classes * ((lp.q * lp.q + 63) / 64);
        }
        lists[p + 1] = lists[p] + lb;
    }
    g.thr_blocks = thr[n];
    g.total_bins = bins[n];
    g.list_blocks = lists[n];
    g.sort_blocks = (max_samples + fg::kBlock - 1) / fg::kBlock;
    if (g.thr_blocks > 0x7fffffffLL || g.list_blocks > 0x7fffffffLL) return set_error(b->ctx, FG_EINVAL, "batch too large");
    // planes by radius distribution, each with the first draw chunk of its
    // bins in that sequence
    std::vector<int> kp[2];
    std::vector<long long> ko[2];
    for (int k = 0; k < 2; ++k) ko[k].push_back(0);
    for (int p = 0; p < n; ++p) {
        const int k = b->kind[p] == 0 ? 0 : 1;
        kp[k].push_back(p);
        ko[k].push_back(ko[k].back() + (bins[p + 1] - bins[p] + fg::kDrawChunk - 1) / fg::kDrawChunk);
    }
    put(thr.data(), thr.size() * 8, reinterpret_cast<const void **>(&g.thr_off));
    put(bins.data(), bins.size() * 8, reinterpret_cast<const void **>(&g.bin_off));
    put(T.data(), T.size() * sizeof(double *), reinterpret_cast<const void **>(&g.T));
    for (int k = 0; k < 2; ++k) {
        g.n_kind[k] = (int)kp[k].size();
        g.kind_chunks[k] = ko[k].back();
        put(kp[k].data(), kp[k].size() * 4, reinterpret_cast<const void **>(&g.kind_plane[k]));
        put(ko[k].data(), ko[k].size() * 8, reinterpret_cast<const void **>(&g.kind_off[k]));
    }
    put(plan.data(), plan.size() * sizeof(int4), reinterpret_cast<const void **>(&g.list_plan));
    put(fill.data(), fill.size() * sizeof(uint32_t *), reinterpret_cast<const void **>(&g.list_fill));
    put(lists.data(), lists.size() * 8, reinterpret_cast<const void **>(&g.list_off));
    int rc = batch_alloc(b, (void **)&b->gen_mem, host.size(), what);
    if (rc != FG_OK) return rc;
    for (const auto &f : fixups) *f.second = b->gen_mem + f.first;
    g.planes = b->d_planes;
    g.cnt = b->cnt;
    g.mask = b->mask;
    hipError_t e = hipMemcpyAsync(b->gen_mem, host.data(), host.size(), hipMemcpyHostToDevice, s);
    // the count pass reads the plane records (their grain fields are set
    // after it and uploaded again)
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_planes, b->planes.data(), sizeof(fg::BatchPlane) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? FG_OK : hip_error(b->ctx, e, what);
}

int batch_create(fg_ctx *ctx, const fg_plane_desc *descs, int n, const float *const *lams,
                 const float *const *offs, float *const *outs, int capture, hipStream_t s, fg_batch *b)
{
    const char *what = "Batch renderer";
    b->ctx = ctx;
    b->n = n;
    int rc;
    for (int p = 0; p < n; ++p) {
        if ((rc = validate(ctx, &descs[p])) != FG_OK) return rc;
        if (!lams[p] || !offs[p] || !outs[p]) return set_error(ctx, FG_EINVAL, "null buffer in batch");
        if (descs[p].n_samples > kHbmMaxSamples)
            return set_error(ctx, FG_EINVAL, "batch planes need n_samples <= 2048");
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if ((rc = batch_alloc(b, (void **)&b->ranges, sizeof(float4) * n, what)) != FG_OK) return rc;
    if ((rc = batch_alloc(b, (void **)&b->status, sizeof(int), what)) != FG_OK) return rc;
    for (int p = 0; p < n && e == hipSuccess; ++p)
        e = fg::launch_offsets_range(reinterpret_cast<const float2 *>(offs[p]), (int)descs[p].n_samples,
                                     b->ranges + p, s);
    std::vector<float4> r(n);
    if (e == hipSuccess) e = hipMemcpyAsync(r.data(), b->ranges, sizeof(float4) * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, what);

    // per plane: geometry, launch shape, arguments, buffers
    b->planes.resize(n);
    b->kind.resize(n);
    b->T.resize(n);
    b->bin_off.resize(n);
    b->samples.resize(n);
    b->qb.resize(n);
    b->list_plans.resize(n);
    b->list_mem.resize(n);
    std::vector<int> lds_p(n);
    long long items = 0;
    // one workgroup size for the whole batch: the explicit one, else 512
    b->block = ctx->cfg_explicit && ctx->cfg.pix_block == 256 ? 256 : 512;
    for (int p = 0; p < n; ++p) {
        const fg_plane_desc *d = &descs[p];
        fg::BatchPlane &bp = b->planes[p];
        if (!plane_geometry(d, r[p], &bp.geo))
            return set_error(ctx, FG_EINVAL, "batch plane " + std::to_string(p) + ": cell grid exceeds 65535 cells per side");
        int tw, th;
        if (!hbm_config(ctx, d, &tw, &th, &lds_p[p], 0, b->block))
            return set_error(ctx, FG_EINVAL, "batch plane " + std::to_string(p) + ": no tile fits the LDS");
        if ((rc = pixel_args(ctx, d, lams[p], offs[p], outs[p], tw, th, lds_p[p], b->ranges + p, b->status,
                             &bp.a)) != FG_OK)
            return rc;
        splat_bounds(bp.a, bp.geo, r[p], &bp.st.eps, &bp.st.safe_r);
        b->kind[p] = (int)d->dist_kind;
        b->bin_off[p] = b->total_bins;
        b->total_bins += (long long)bp.geo.nbx * bp.geo.nby;
        bp.item_base = items;
        items += bp.a.n_tiles;
        b->qb[p] = sample_buckets(d->n_samples);
        if ((rc = batch_alloc(b, (void **)&b->T[p], (size_t)d->in_w * d->in_h * sizeof(double), what)) != FG_OK)
            return rc;
        if ((rc = batch_alloc(b, &b->samples[p], fg::sample_table_bytes((int)d->n_samples, b->qb[p]), what)) != FG_OK)
            return rc;
        fg::sample_table_layout(b->samples[p], (int)d->n_samples, b->qb[p], &bp.st);
        // candidate lists, built with the plane's table on every run
        b->list_plans[p] = list_plan(bp.a, bp.st.eps);
        b->list_mem[p] = nullptr;
        const size_t lb = b->list_plans[p].bytes((int)d->n_samples);
        if (lb > 0 && lb < ((size_t)1 << 32)) {
            if ((rc = batch_alloc(b, (void **)&b->list_mem[p], lb, what)) != FG_OK) return rc;
            const ListPlan &lp = b->list_plans[p];
            bp.st.cl_cell = reinterpret_cast<const uint2 *>(b->list_mem[p] + 256);
            bp.st.cl_yidx = reinterpret_cast<const uint16_t *>(bp.st.cl_cell + (size_t)lp.classes * lp.q * lp.q);
            bp.st.cl_list = bp.st.cl_yidx + (size_t)lp.classes * lp.q * lp.q * fg::kRowBuckets;
            bp.a.cl_q = lp.q;
            bp.a.cl_classes = lp.classes;
            bp.a.cl_inv_dz = (float)(2 * lp.q);
        }
    }
    if (items > 0x7fffffffLL || b->total_bins >= 0xffffffffLL)
        return set_error(ctx, FG_EINVAL, "batch too large");
    b->n_items = (int)items;
    // planes that share the description and the offsets' range share every
    // cell's stream (the same seed and grid)
    b->count_groups.clear();
    for (int p = 0; p < n; ++p) {
        if (!b->count_groups.empty()) {
            auto &g = b->count_groups.back();
            const int q = g.first;
            if (g.second < fg::kFramePlanes && std::memcmp(&descs[q], &descs[p], sizeof(fg_plane_desc)) == 0 &&
                std::memcmp(&r[q], &r[p], sizeof(float4)) == 0) {
                ++g.second;
                continue;
            }
        }
        b->count_groups.push_back({p, 1});
    }
    b->lds = *std::max_element(lds_p.begin(), lds_p.end());
    if ((size_t)b->lds + fg::splat_hbm_static_lds(b->block) > (size_t)ctx->max_lds)
        return set_error(ctx, FG_EINVAL, "batch planes need more LDS than a workgroup has");
    const long long nb = b->total_bins;
    if ((rc = batch_alloc(b, (void **)&b->cnt, (size_t)nb * 4, what)) != FG_OK) return rc;
    if ((rc = batch_alloc(b, (void **)&b->start, (size_t)(nb + 1) * 4, what)) != FG_OK) return rc;
    if ((rc = batch_alloc(b, (void **)&b->partials, fg::scan_partials(nb) * 4, what)) != FG_OK) return rc;
    if ((rc = batch_alloc(b, (void **)&b->mask, (size_t)nb * 32, what)) != FG_OK) return rc;
    if ((rc = batch_alloc(b, (void **)&b->weights, (size_t)b->n_items * 4, what)) != FG_OK) return rc;
    if ((rc = batch_alloc(b, (void **)&b->items, (size_t)b->n_items * sizeof(int2), what)) != FG_OK) return rc;
    if ((rc = batch_alloc(b, (void **)&b->counter, 4, what)) != FG_OK) return rc;
    if ((rc = batch_alloc(b, (void **)&b->total64, 16, what)) != FG_OK) return rc;
    if ((rc = batch_alloc(b, (void **)&b->d_planes, sizeof(fg::BatchPlane) * n, what)) != FG_OK) return rc;

    if ((rc = batch_gen_setup(b, what, s)) != FG_OK) return rc;
    // counting pass over the current inputs sizes the grain array
    e = batch_enqueue_counts(b, s);
    unsigned long long total = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&total, b->total64, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if (total > kMaxGrains)
        return set_error(ctx, FG_ECAPACITY,
                         std::string(what) + ": the batch has " + std::to_string(total) + " grains, more than the " +
                             std::to_string(kMaxGrains) + " one batch addresses; split it");
    const double want = (double)total * 1.25 + 65536.0;
    b->cap = (uint32_t)std::min(want, (double)kMaxGrains);
    if ((rc = batch_alloc(b, (void **)&b->grains, (size_t)b->cap * sizeof(float4), what)) != FG_OK) return rc;
    for (int p = 0; p < n; ++p) {
        fg::BatchPlane &bp = b->planes[p];
        bp.a.lds_bytes = b->lds;
        bp.a.grain_total = b->total64;
        bp.a.grain_cap = b->cap;
        bp.bin_start = b->start + b->bin_off[p];
        bp.grains = b->grains;
    }
    b->gen.grains = b->grains;
    e = hipMemcpyAsync(b->d_planes, b->planes.data(), sizeof(fg::BatchPlane) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    b->grid = (int)std::max<long long>(
        1, std::min<long long>(std::max(b->n_items, 1),
                               (long long)ctx->num_cus * fg::splat_batch_blocks_per_cu(b->block, b->lds)));
    if (capture) {
        e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            const hipError_t e2 = batch_enqueue(b, b->capture_stream);
            e = hipStreamEndCapture(b->capture_stream, &b->graph);
            if (e2 != hipSuccess) e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return hip_error(ctx, e, what);
    }
    return FG_OK;
}

}  // namespace

extern "C" {

int fg_ctx_create(int device_ordinal, fg_ctx **out)
{
    if (!out) return FG_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return FG_ENODEV;
    }
    if (device_ordinal < 0 || device_ordinal >= count) return FG_ENODEV;
    fg_ctx *ctx = new fg_ctx();
    ctx->device = device_ordinal;
    ctx->table.assign(kBuiltinLaunchTable, kBuiltinLaunchTable + kBuiltinLaunchCount);
    hipError_t e = hipSetDevice(device_ordinal);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device_ordinal);
    if (e == hipSuccess) {
        ctx->max_lds = (int)prop.sharedMemPerBlock;
        ctx->num_cus = prop.multiProcessorCount;
        e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    }
    if (e == hipSuccess) e = hipMalloc(&ctx->d_status, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&ctx->d_stats, 4 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&ctx->d_sum64, 2 * sizeof(unsigned long long));
    // this device's copy of the allocation-free entry point's scratch (the
    // device is current)
    if (e == hipSuccess) e = hipGetSymbolAddress(reinterpret_cast<void **>(&ctx->noalloc_scratch), HIP_SYMBOL(scratch));
    if (e == hipSuccess) e = hipMalloc(&ctx->d_range, sizeof(float4));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->scratch_done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemset(ctx->d_status, 0, sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->h_cancel, 64, hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess) {
        *ctx->h_cancel = 0;
        e = hipHostGetDevicePointer((void **)&ctx->d_cancel, ctx->h_cancel, 0);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        fg_ctx_destroy(ctx);
        return FG_EHIP;
    }
    *out = ctx;
    return FG_OK;
}

void fg_ctx_destroy(fg_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_buffers(ctx);
    (void)hipFree(ctx->d_status);
    (void)hipFree(ctx->d_stats);
    (void)hipFree(ctx->d_sum64);
    (void)hipFree(ctx->d_range);
    if (ctx->h_cancel) (void)hipHostFree(ctx->h_cancel);
    if (ctx->scratch_done) (void)hipEventDestroy(ctx->scratch_done);
    for (hipEvent_t ev : ctx->tev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *fg_last_error(const fg_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

int fg_cancel(fg_ctx *ctx, int on)
{
    if (!ctx) return FG_EINVAL;
    set_cancel(ctx, on ? 1 : 0);
    return FG_OK;
}

int fg_render_pixelwise(fg_ctx *ctx, const fg_plane_desc *desc, const float *lambda,
                        const float *offsets_input_xy, float *out, fg_cancel_fn cancel,
                        void *cancel_user)
{
    return render_host(ctx, desc, lambda, offsets_input_xy, out, cancel, cancel_user, do_pixelwise,
                       "Pixel renderer");
}

int fg_render_grainwise(fg_ctx *ctx, const fg_plane_desc *desc, const float *lambda,
                        const float *offsets_xy, float *out, fg_cancel_fn cancel, void *cancel_user)
{
    return render_host(ctx, desc, lambda, offsets_xy, out, cancel, cancel_user, do_grainwise,
                       "Grain renderer");
}

int fg_render_pixelwise_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                               const float *d_offsets_input_xy, float *d_out, void *hip_stream)
{
    int rc = validate(ctx, desc);
    if (rc != FG_OK) return rc;
    if (!d_lambda || !d_offsets_input_xy || !d_out) return set_error(ctx, FG_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, "Pixel renderer");
    return do_pixelwise(ctx, desc, d_lambda, d_offsets_input_xy, d_out,
                        reinterpret_cast<hipStream_t>(hip_stream), false);
}

int fg_render_pixelwise_planes_device(fg_ctx *ctx, const fg_plane_desc *desc, int n_planes,
                                      const float *const *d_lambdas, const float *d_offsets_input_xy,
                                      float *const *d_outs, void *hip_stream)
{
    int rc = validate(ctx, desc);
    if (rc != FG_OK) return rc;
    if (n_planes < 1 || n_planes > 3) return set_error(ctx, FG_EINVAL, "n_planes must be 1, 2 or 3");
    if (!d_lambdas || !d_outs || !d_offsets_input_xy) return set_error(ctx, FG_EINVAL, "null buffer");
    for (int i = 0; i < n_planes; ++i)
        if (!d_lambdas[i] || !d_outs[i]) return set_error(ctx, FG_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, "Pixel renderer");
    return do_pixelwise_planes(ctx, desc, n_planes, d_lambdas, d_offsets_input_xy, d_outs,
                               reinterpret_cast<hipStream_t>(hip_stream));
}

int fg_render_pixelwise_device_noalloc(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                                       const float *d_offsets_input_xy, float *d_out, void *hip_stream)
{
    const char *what = "Pixel renderer (allocation-free)";
    int rc = validate(ctx, desc);
    if (rc != FG_OK) return rc;
    if (!d_lambda || !d_offsets_input_xy || !d_out) return set_error(ctx, FG_EINVAL, "null buffer");
    if (desc->n_samples > kHbmMaxSamples)
        return set_error(ctx, FG_EINVAL, std::string(what) + ": n_samples must be <= 2048");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if (!ctx->noalloc_scratch) return set_error(ctx, FG_EHIP, std::string(what) + ": no scratch address");
    hipStream_t s = reinterpret_cast<hipStream_t>(hip_stream);
    float4 range{};
    e = fg::launch_offsets_range(reinterpret_cast<const float2 *>(d_offsets_input_xy), (int)desc->n_samples,
                                 ctx->d_range, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&range, ctx->d_range, sizeof range, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    fg::PixGeom geom{};
    int tw, th, lds;
    if (!plane_geometry(desc, range, &geom) || !hbm_config(ctx, desc, &tw, &th, &lds, 4LL * ctx->num_cus))
        return set_error(ctx, FG_EINVAL, std::string(what) + ": the plane's cell grid does not fit the HBM path");
    fg::PixelArgs a{};
    if ((rc = pixel_args(ctx, desc, d_lambda, d_offsets_input_xy, d_out, tw, th, lds, ctx->d_range, ctx->d_status,
                         &a)) != FG_OK)
        return rc;
    ScratchSource src;
    src.arena = ctx->noalloc_scratch;
    src.cap = kNoAllocScratchBytes;
    return run_pixel_hbm(ctx, a, geom, range, (int)desc->dist_kind, splat_hbm_block(ctx, desc), s, src);
}

int fg_render_pixelwise_adaptive_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                                        const float *d_offsets_input_xy, const uint16_t *d_budget, float *d_out,
                                        void *hip_stream)
{
    int rc = validate(ctx, desc);
    if (rc != FG_OK) return rc;
    if (!d_lambda || !d_offsets_input_xy || !d_budget || !d_out) return set_error(ctx, FG_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, "Pixel renderer");
    return do_pixelwise_budget(ctx, desc, d_lambda, d_offsets_input_xy, d_out,
                               reinterpret_cast<hipStream_t>(hip_stream), false, d_budget);
}

int fg_render_grainwise_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                               const float *d_offsets_xy, float *d_out, void *hip_stream)
{
    int rc = validate(ctx, desc);
    if (rc != FG_OK) return rc;
    if (!d_lambda || !d_offsets_xy || !d_out) return set_error(ctx, FG_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, "Grain renderer");
    return do_grainwise(ctx, desc, d_lambda, d_offsets_xy, d_out,
                        reinterpret_cast<hipStream_t>(hip_stream), false);
}

int fg_stream_status(fg_ctx *ctx, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, "status");
    return read_status(ctx, reinterpret_cast<hipStream_t>(hip_stream), "render");
}

int fg_pixelwise_stats_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                              const float *d_offsets_input_xy, float *d_out, uint64_t counters[3],
                              void *hip_stream)
{
    return stats_device(ctx, desc, d_lambda, d_offsets_input_xy, d_out, counters, 3, hip_stream,
                        do_pixelwise, "Pixel renderer");
}

int fg_grainwise_stats_device(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                              const float *d_offsets_xy, float *d_out, uint64_t counters[4],
                              void *hip_stream)
{
    return stats_device(ctx, desc, d_lambda, d_offsets_xy, d_out, counters, 4, hip_stream,
                        do_grainwise, "Grain renderer");
}

int fg_set_launch_config(fg_ctx *ctx, const fg_launch_config *cfg)
{
    if (!ctx) return FG_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (cfg) {
        if (cfg->pix_tile_w < 0 || cfg->pix_tile_h < 0 ||
            cfg->pix_tile_w > 256 || cfg->pix_tile_h > 256 || cfg->grain_tile_w < 0 ||
            cfg->grain_tile_h < 0 || cfg->grain_grains_per_thread < 0 ||
            cfg->grain_grains_per_thread > 64 || cfg->pix_lds_bytes < 0 || cfg->grain_lds_bytes < 0 ||
            cfg->pix_variant < 0 || cfg->pix_variant > 3 ||
            (cfg->pix_block != 0 && cfg->pix_block != 256 && cfg->pix_block != 512) || cfg->grain_variant < 0 ||
            cfg->grain_variant > 2 || (cfg->grain_block != 0 && cfg->grain_block != 256 && cfg->grain_block != 512))
            return set_error(ctx, FG_EINVAL, "invalid launch configuration");
        ctx->cfg = *cfg;
        const fg_launch_config zero{};
        ctx->cfg_explicit = std::memcmp(cfg, &zero, sizeof zero) != 0;
    } else {
        ctx->cfg = fg_launch_config{};
        ctx->cfg_explicit = false;
    }
    return FG_OK;
}

int fg_launch_bucket(const fg_plane_desc *desc, int algo, fg_launch_key *out)
{
    if (!desc || !out || (algo != 0 && algo != 1)) return FG_EINVAL;
    *out = launch_key(desc, algo);
    return FG_OK;
}

int fg_set_launch_table(fg_ctx *ctx, const fg_launch_entry *entries, int n)
{
    if (!ctx) return FG_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (n < 0) {
        ctx->table.assign(kBuiltinLaunchTable, kBuiltinLaunchTable + kBuiltinLaunchCount);
        return FG_OK;
    }
    if (n > 0 && !entries) return set_error(ctx, FG_EINVAL, "null launch table");
    for (int i = 0; i < n; ++i) {
        const fg_launch_config &c = entries[i].cfg;
        if (entries[i].key.algo > 1 || entries[i].key.dist_kind > 1 || c.pix_tile_w < 0 || c.pix_tile_h < 0 ||
            c.pix_tile_w > 256 || c.pix_tile_h > 256 || c.grain_tile_w < 0 || c.grain_tile_h < 0 ||
            c.pix_variant < 0 || c.pix_variant > 3 || (c.pix_block != 0 && c.pix_block != 256 && c.pix_block != 512) ||
            c.grain_variant < 0 || c.grain_variant > 2 ||
            (c.grain_block != 0 && c.grain_block != 256 && c.grain_block != 512) || c.pix_lds_bytes < 0 ||
            c.grain_lds_bytes < 0 || c.grain_grains_per_thread < 0 || c.grain_grains_per_thread > 64)
            return set_error(ctx, FG_EINVAL, "invalid launch table entry " + std::to_string(i));
    }
    ctx->table.assign(entries, entries + n);
    return FG_OK;
}

int fg_get_launch_table(fg_ctx *ctx, fg_launch_entry *out, int cap)
{
    if (!ctx) return FG_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const int n = (int)ctx->table.size();
    for (int i = 0; i < n && i < cap && out; ++i) out[i] = ctx->table[i];
    return n;
}

int fg_get_launch_config(fg_ctx *ctx, const fg_plane_desc *desc, int algo, fg_launch_config *out)
{
    int rc = validate(ctx, desc);
    if (rc != FG_OK) return rc;
    if (!out) return set_error(ctx, FG_EINVAL, "null output");
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = fg_launch_config{};
    if (algo == 1) {
        out->grain_variant = grain_variant(ctx, desc);
        out->grain_block = grain_block(ctx, desc);
        if (out->grain_variant == 2 &&
            !sorted_grain_config(ctx, desc, &out->grain_tile_w, &out->grain_tile_h, &out->grain_lds_bytes))
            out->grain_variant = 1;
        if (out->grain_variant == 1)
            grain_config(ctx, desc, &out->grain_tile_w, &out->grain_tile_h, &out->grain_lds_bytes,
                         &out->grain_grains_per_thread);
    } else {
        out->pix_variant = pixel_variant(ctx, desc);
        out->pix_block = splat_hbm_block(ctx, desc);
        if (out->pix_variant == 3 &&
            !hbm_config(ctx, desc, &out->pix_tile_w, &out->pix_tile_h, &out->pix_lds_bytes, 4LL * ctx->num_cus))
            out->pix_variant = 2;
        if (out->pix_variant != 3)
            pixel_config(ctx, desc, out->pix_variant, &out->pix_tile_w, &out->pix_tile_h,
                         &out->pix_lds_bytes);
    }
    return FG_OK;
}

int fg_set_timing(fg_ctx *ctx, int enable)
{
    if (!ctx) return FG_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->timing = enable != 0;
    ctx->timed = false;
    return FG_OK;
}

int fg_get_timings(fg_ctx *ctx, float ms[3])
{
    if (!ctx || !ms) return FG_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->timed) return set_error(ctx, FG_EINVAL, "no timed render on this context");
    hipError_t e = hipEventSynchronize(ctx->tev[3]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[0], ctx->tev[0], ctx->tev[3]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[1], ctx->tev[0], ctx->tev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[2], ctx->tev[2], ctx->tev[3]);
    if (e != hipSuccess) return hip_error(ctx, e, "timing");
    return FG_OK;
}

int fg_batch_create(fg_ctx *ctx, const fg_plane_desc *descs, int n_planes, const float *const *d_lambdas,
                    const float *const *d_offsets_input_xy, float *const *d_outs, int capture_graph,
                    void *hip_stream, fg_batch **out)
{
    if (!ctx || !out) return FG_EINVAL;
    *out = nullptr;
    if (n_planes <= 0 || !descs || !d_lambdas || !d_offsets_input_xy || !d_outs)
        return set_error(ctx, FG_EINVAL, "empty batch");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = hip_stream ? reinterpret_cast<hipStream_t>(hip_stream) : ctx->stream;
    fg_batch *b = new fg_batch();
    const int rc = batch_create(ctx, descs, n_planes, d_lambdas, d_offsets_input_xy, d_outs, capture_graph, s, b);
    if (rc != FG_OK) {
        (void)hipStreamSynchronize(s);
        batch_free(b);
        delete b;
        return rc;
    }
    *out = b;
    return FG_OK;
}

int fg_batch_run(fg_batch *b, void *hip_stream)
{
    if (!b) return FG_EINVAL;
    fg_ctx *ctx = b->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = hip_stream ? reinterpret_cast<hipStream_t>(hip_stream) : ctx->stream;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = b->exec ? hipGraphLaunch(b->exec, s) : batch_enqueue(b, s);
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "Batch renderer");
}

int fg_batch_status(fg_batch *b, void *hip_stream)
{
    if (!b) return FG_EINVAL;
    fg_ctx *ctx = b->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = hip_stream ? reinterpret_cast<hipStream_t>(hip_stream) : ctx->stream;
    int status = 0;
    hipError_t e = hipMemcpyAsync(&status, b->status, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, "Batch renderer");
    return status_error(ctx, status, "Batch renderer");
}

int fg_batch_info(const fg_batch *b, int *n_items, int *grid_blocks, int *graph_captured)
{
    if (!b) return FG_EINVAL;
    if (n_items) *n_items = b->n_items;
    if (grid_blocks) *grid_blocks = b->grid;
    if (graph_captured) *graph_captured = b->exec ? 1 : 0;
    return FG_OK;
}

void fg_batch_destroy(fg_batch *b)
{
    if (!b) return;
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    batch_free(b);
    delete b;
}

int fg_mlp_features_device(fg_ctx *ctx, const float *d_u, uint32_t in_w, uint32_t in_h, uint32_t out_w,
                           uint32_t out_h, float zoom, float radius_mean, float radius_stddev, float sigma,
                           int channel, uint16_t *d_features, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (!d_u || !d_features) return set_error(ctx, FG_EINVAL, "null buffer");
    if (in_w == 0 || in_h == 0 || out_w == 0 || out_h == 0 || !(zoom > 0.0f) || channel < 0 || channel > 2)
        return set_error(ctx, FG_EINVAL, "invalid feature request");
    const hipError_t e = fg::launch_features(d_u, (int)in_w, (int)in_h, (int)out_w, (int)out_h, zoom, radius_mean,
                                             radius_stddev, sigma, channel, d_features,
                                             reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "features");
}

int fg_mlp_forward_device(fg_ctx *ctx, const float *d_params, const uint16_t *d_features, int64_t n, float *d_out,
                          void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (!d_params || !d_features || !d_out || n < 0) return set_error(ctx, FG_EINVAL, "invalid MLP arguments");
    if (n == 0) return FG_OK;
    const hipError_t e = fg::launch_mlp_forward(d_params, d_features, n, d_out, reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "MLP forward");
}

int fg_mlp_grad_device(fg_ctx *ctx, const float *d_params, const uint16_t *d_features, const float *d_target,
                       int64_t n, int64_t n_total, int loss_kind, float *d_grad, float *d_loss, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (!d_params || !d_features || !d_target || !d_grad || !d_loss || n < 0 || n_total <= 0 ||
        (loss_kind != FG_LOSS_SQUARED && loss_kind != FG_LOSS_NLL))
        return set_error(ctx, FG_EINVAL, "invalid MLP arguments");
    if (n == 0) return FG_OK;
    const hipError_t e = fg::launch_mlp_grad(d_params, d_features, d_target, n, 1.0f / (float)n_total, loss_kind,
                                             d_grad, d_loss, reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "MLP gradient");
}

int fg_render_frame_device(fg_ctx *ctx, const fg_frame_desc *fd, const float *d_rgb, const float *d_offsets_xy,
                           uint8_t *d_rgb8, float *d_planes, void *hip_stream)
{
    int rc = validate_frame(ctx, fd);
    if (rc != FG_OK) return rc;
    if (!d_rgb || !d_offsets_xy || !d_rgb8) return set_error(ctx, FG_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, "Frame renderer");
    return frame_device(ctx, fd, d_rgb, d_offsets_xy, d_rgb8, d_planes, reinterpret_cast<hipStream_t>(hip_stream),
                        nullptr, nullptr);
}

int fg_render_frame(fg_ctx *ctx, const fg_frame_desc *fd, const float *rgb, const float *offsets_xy, uint8_t *rgb8,
                    float *planes, fg_cancel_fn cancel, void *cancel_user)
{
    const char *what = "Frame renderer";
    int rc = validate_frame(ctx, fd);
    if (rc != FG_OK) return rc;
    if (!rgb || !offsets_xy || !rgb8) return set_error(ctx, FG_EINVAL, "null buffer");
    const fg_plane_desc &d = fd->plane;
    if (!finite_pairs(offsets_xy, d.n_samples)) return set_error(ctx, FG_EINVAL, "offsets must be finite");
    if (cancel && cancel(cancel_user)) return set_error(ctx, FG_ECANCELLED, "cancelled");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    const size_t nin = (size_t)d.in_w * d.in_h, nout = (size_t)d.out_w * d.out_h;
    const int nr = fd->color_mode == FG_COLOR_RGB ? 3 : 1;
    const size_t rgb_b = align256(nin * 12), offs_b = align256((size_t)d.n_samples * 8), out_b = align256(nout * 3);
    const size_t planes_b = planes ? align256((size_t)nr * nout * 4) : 0;
    if ((rc = ensure_bytes(ctx, ctx->frame_io, rgb_b + offs_b + out_b + planes_b, what)) != FG_OK) return rc;
    unsigned char *base = static_cast<unsigned char *>(ctx->frame_io.p);
    float *d_rgb = reinterpret_cast<float *>(base);
    float *d_offs = reinterpret_cast<float *>(base + rgb_b);
    uint8_t *d_rgb8 = base + rgb_b + offs_b;
    float *d_planes = planes ? reinterpret_cast<float *>(base + rgb_b + offs_b + out_b) : nullptr;
    hipStream_t s = ctx->stream;
    e = hipMemcpyAsync(d_rgb, rgb, nin * 12, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_offs, offsets_xy, (size_t)d.n_samples * 8, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    set_cancel(ctx, 0);
    rc = frame_device(ctx, fd, d_rgb, d_offs, d_rgb8, d_planes, s, cancel, cancel_user);
    if (rc != FG_OK) {
        (void)hipStreamSynchronize(s);
        (void)read_status(ctx, s, what);
        set_cancel(ctx, 0);
        return rc;
    }
    e = wait_polling(ctx, s, cancel, cancel_user);
    if (e == hipSuccess) e = hipMemcpyAsync(rgb8, d_rgb8, nout * 3, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && planes)
        e = hipMemcpyAsync(planes, d_planes, (size_t)nr * nout * 4, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    rc = read_status(ctx, s, what);
    set_cancel(ctx, 0);
    return rc;
}

int fg_lambda_device(fg_ctx *ctx, const float *d_plane, uint32_t w, uint32_t h, float inv_e_pi_r2, float *d_norm,
                     float *d_lambda, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (!d_plane || !d_lambda) return set_error(ctx, FG_EINVAL, "null buffer");
    if (!std::isfinite(inv_e_pi_r2)) return set_error(ctx, FG_EINVAL, "inv_e_pi_r2 must be finite");
    if ((size_t)w * h == 0) return FG_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, "Lambda pre-pass");
    int rc = scratch_acquire(ctx, "Lambda pre-pass");
    if (rc != FG_OK) return rc;
    if ((rc = ensure_bytes(ctx, ctx->frame, 256, "Lambda pre-pass")) != FG_OK) return rc;
    // the max word: the first word of the frame scratch
    const hipStream_t s = reinterpret_cast<hipStream_t>(hip_stream);
    e = fg::launch_lambda(d_plane, (long long)w * h, inv_e_pi_r2, static_cast<unsigned *>(ctx->frame.p), d_norm,
                          d_lambda, s);
    if (e != hipSuccess) return hip_error(ctx, e, "Lambda pre-pass");
    return scratch_release(ctx, s, "Lambda pre-pass");
}

int fg_color_split_device(fg_ctx *ctx, int color_mode, const float *d_rgb, uint32_t w, uint32_t h, float *d_p0,
                          float *d_p1, float *d_p2, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (color_mode != FG_COLOR_LUMA && color_mode != FG_COLOR_RGB)
        return set_error(ctx, FG_EINVAL, "unknown colour mode");
    if (!d_rgb || !d_p0 || !d_p1 || !d_p2) return set_error(ctx, FG_EINVAL, "null buffer");
    if ((size_t)w * h == 0) return FG_OK;
    const hipError_t e = fg::launch_split(color_mode, d_rgb, (long long)w * h, d_p0, d_p1, d_p2,
                                          reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "Colour split");
}

int fg_color_merge_device(fg_ctx *ctx, int color_mode, const float *d_p0, const float *d_p1, const float *d_p2,
                          uint32_t w, uint32_t h, uint32_t cw, uint32_t ch, uint8_t *d_rgb8, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (color_mode != FG_COLOR_LUMA && color_mode != FG_COLOR_RGB)
        return set_error(ctx, FG_EINVAL, "unknown colour mode");
    if (!d_p0 || !d_p1 || !d_p2 || !d_rgb8) return set_error(ctx, FG_EINVAL, "null buffer");
    if (w > (1u << 24) || h > (1u << 24) || cw > (1u << 24) || ch > (1u << 24))
        return set_error(ctx, FG_EINVAL, "image dimensions exceed 2^24");
    if (color_mode == FG_COLOR_LUMA && (cw == 0 || ch == 0) && (size_t)w * h > 0)
        return set_error(ctx, FG_EINVAL, "chroma planes are empty");
    if ((size_t)w * h == 0) return FG_OK;
    const hipError_t e = fg::launch_merge(color_mode, d_p0, d_p1, d_p2, (int)w, (int)h,
                                          color_mode == FG_COLOR_RGB ? (int)w : (int)cw,
                                          color_mode == FG_COLOR_RGB ? (int)h : (int)ch, d_rgb8,
                                          reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "Colour merge");
}

int fg_math_device(fg_ctx *ctx, int fn, const double *d_in, int64_t n, double *d_out, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (fn < 0 || fn > fg::kMathExpF32) return set_error(ctx, FG_EINVAL, "unknown math function");
    if (n < 0 || (n > 0 && (!d_in || !d_out))) return set_error(ctx, FG_EINVAL, "null buffer");
    if (n == 0) return FG_OK;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = fg::launch_math_probe(fn, d_in, n, d_out, reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "math");
}

int fg_log_f32_device(fg_ctx *ctx, const float *d_in, int64_t n, float *d_out, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (n < 0 || (n > 0 && (!d_in || !d_out))) return set_error(ctx, FG_EINVAL, "invalid log arguments");
    if (n == 0) return FG_OK;
    const hipError_t e = fg::launch_log_f32(d_in, n, d_out, reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "Log");
}

int fg_mlp_budget_device(fg_ctx *ctx, const float *d_pred_log_s, const uint16_t *d_features, int64_t n, float c,
                         int n_min, int n_max, uint16_t *d_budget, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if ((!d_pred_log_s && !d_features) || !d_budget || n < 0 || n_min < 1 || n_max < n_min || n_max > 65535 ||
        !(c >= 0.0f))
        return set_error(ctx, FG_EINVAL, "invalid budget arguments");
    if (n == 0) return FG_OK;
    const hipError_t e = fg::launch_budget(d_pred_log_s, d_features, n, c, nullptr, n_min, n_max, d_budget,
                                           reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "budget");
}

int fg_mlp_budget_scale_device(fg_ctx *ctx, const float *d_pred_log_s, const uint16_t *d_features, int64_t n,
                               const float *d_c, int n_min, int n_max, uint16_t *d_budget, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if ((!d_pred_log_s && !d_features) || !d_c || !d_budget || n < 0 || n_min < 1 || n_max < n_min ||
        n_max > 65535)
        return set_error(ctx, FG_EINVAL, "invalid budget arguments");
    if (n == 0) return FG_OK;
    const hipError_t e = fg::launch_budget(d_pred_log_s, d_features, n, 0.0f, d_c, n_min, n_max, d_budget,
                                           reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "budget");
}

int fg_budget_scale_device(fg_ctx *ctx, const float *d_pred_log_s, int64_t n, float s_mult, double target_mse,
                           int n_min, int n_max, int n_ref, float *d_scale, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (!d_pred_log_s || !d_scale || n < 1 || n_min < 1 || n_max < n_min || n_ref < 1 || !(s_mult > 0.0f) ||
        !(target_mse >= 0.0))
        return set_error(ctx, FG_EINVAL, "invalid budget scale arguments");
    const char *what = "budget scale";
    const int rc = ensure_bytes(ctx, ctx->scale_state, fg::budget_scale_state_bytes(), what);
    if (rc != FG_OK) return rc;
    const hipError_t e = fg::launch_budget_scale(d_pred_log_s, n, s_mult, target_mse, n_min, n_max, n_ref, 3,
                                                 (long long)1 << 18, ctx->scale_state.p, d_scale,
                                                 reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, what);
}

int fg_mlp_adam_device(fg_ctx *ctx, float *d_params, const float *d_grad, float *d_m, float *d_v, float lr,
                       float beta1, float beta2, float eps, int step, void *hip_stream)
{
    if (!ctx) return FG_EINVAL;
    if (!d_params || !d_grad || !d_m || !d_v || step < 1) return set_error(ctx, FG_EINVAL, "invalid Adam arguments");
    const float bc1 = 1.0f - std::pow(beta1, (float)step), bc2 = 1.0f - std::pow(beta2, (float)step);
    const hipError_t e = fg::launch_mlp_adam(d_params, d_grad, d_m, d_v, lr, beta1, beta2, eps, bc1, bc2,
                                             reinterpret_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? FG_OK : hip_error(ctx, e, "Adam step");
}

uint64_t fg_max_grains(void) { return kMaxGrains; }

int fg_pixelwise_grains(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda,
                        const float *d_offsets_input_xy, float *grains_xyrc, uint64_t capacity, uint64_t *n_grains,
                        int32_t origin[2], uint32_t size[2])
{
    const char *what = "Pixel renderer";
    int rc = validate(ctx, desc);
    if (rc != FG_OK) return rc;
    if (!d_lambda || !d_offsets_input_xy || !n_grains || !origin || !size || (capacity && !grains_xyrc))
        return set_error(ctx, FG_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    hipStream_t s = ctx->stream;
    float4 range{};
    e = fg::launch_offsets_range(reinterpret_cast<const float2 *>(d_offsets_input_xy), (int)desc->n_samples,
                                 ctx->d_range, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&range, ctx->d_range, sizeof range, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    fg::PixGeom geom{};
    int tw, th, lds;
    if (!plane_geometry(desc, range, &geom) || !hbm_config(ctx, desc, &tw, &th, &lds, 0))
        return set_error(ctx, FG_EINVAL, "the plane's cell grid does not fit the HBM generation");
    fg::PixelArgs a{};
    if ((rc = pixel_args(ctx, desc, d_lambda, d_offsets_input_xy, nullptr, tw, th, lds, ctx->d_range, ctx->d_status,
                         &a)) != FG_OK)
        return rc;
    if ((rc = scratch_acquire(ctx, what)) != FG_OK) return rc;
    ScratchSource src;
    HbmGen gen;
    if ((rc = hbm_generate(ctx, src, a, geom, (int)desc->dist_kind, s, &gen)) != FG_OK) return rc;
    *n_grains = gen.total;
    origin[0] = geom.Ix0;
    origin[1] = geom.Jy0;
    size[0] = (uint32_t)geom.W;
    size[1] = (uint32_t)geom.H;
    if (gen.total <= capacity && gen.total > 0)
        e = hipMemcpyAsync(grains_xyrc, gen.grains, (size_t)gen.total * sizeof(float4), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    if ((rc = scratch_release(ctx, s, what)) != FG_OK) return rc;
    rc = read_status(ctx, s, what);
    if (rc == FG_OK && gen.total > capacity)
        return set_error(ctx, FG_ECAPACITY, "grain buffer holds fewer records than the plane has");
    return rc;
}

int fg_grainwise_pixel_state(fg_ctx *ctx, const fg_plane_desc *desc, const float *d_lambda, uint32_t *q_out,
                             uint64_t *state_out)
{
    const char *what = "Grain renderer";
    int rc = validate(ctx, desc);
    if (rc != FG_OK) return rc;
    if (!d_lambda || !q_out || !state_out) return set_error(ctx, FG_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    hipStream_t s = ctx->stream;
    if ((rc = scratch_acquire(ctx, what)) != FG_OK) return rc;
    const size_t npx = (size_t)desc->in_w * desc->in_h;
    if ((rc = ensure_bytes(ctx, ctx->pix_state, npx * 36 + 32, what)) != FG_OK) return rc;
    unsigned char *ps = static_cast<unsigned char *>(ctx->pix_state.p);
    fg::GrainArgs a{};
    a.lambda = d_lambda;
    a.status = ctx->d_status;
    a.seed = desc->seed;
    a.in_w = (int)desc->in_w;
    a.in_h = (int)desc->in_h;
    a.pix_s01 = reinterpret_cast<ulonglong2 *>(ps);
    a.pix_s23 = reinterpret_cast<ulonglong2 *>(ps + npx * 16);
    a.pix_q = reinterpret_cast<uint32_t *>(ps + npx * 32);
    std::vector<uint64_t> s01(2 * npx), s23(2 * npx);
    e = fg::launch_pixel_state(a, s);
    if (e == hipSuccess) e = hipMemcpyAsync(q_out, a.pix_q, npx * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(s01.data(), a.pix_s01, npx * 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(s23.data(), a.pix_s23, npx * 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, what);
    for (size_t p = 0; p < npx; ++p) {
        state_out[4 * p] = s01[2 * p];
        state_out[4 * p + 1] = s01[2 * p + 1];
        state_out[4 * p + 2] = s23[2 * p];
        state_out[4 * p + 3] = s23[2 * p + 1];
    }
    if ((rc = scratch_release(ctx, s, what)) != FG_OK) return rc;
    return read_status(ctx, s, what);
}

const char *fg_version(void) { return "film_grain_amd 0.1.0 (gfx950)"; }

}  // extern "C"
