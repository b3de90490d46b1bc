// fg_pixelwise.hip -- gfx950 pixel-wise Boolean-model integrator.
//
// Computes exactly what render_pixelwise / evaluate_indicator compute
// (src/pixelwise.rs:11-106): for every output pixel and Monte-Carlo offset k,
// the point (xg, yg) = ((x+0.5)/zoom - off_k.x, (y+0.5)/zoom - off_k.y) is
// covered iff some grain of some cell (i, j) in the reference's window
// [floor((xg-rm)/delta), floor((xg+rm)/delta)] x [...] contains it; the pixel
// value is hits * (1/N).
//
// The reference regenerates a cell's grains (cell hash, Xoshiro seed, Poisson
// count, centres, radii) at every (pixel, sample, cell) visit.  A cell's grains
// depend only on (seed, i, j, lambda(cell)), so here each workgroup owns an
// output tile, generates every cell of the tile's neighbourhood once into LDS,
// and the coverage tests become pure FP32 work:
//   phase 1  cells in column-major order (i outer, j inner, the reference's
//            loop order), one cell per thread per chunk; Poisson counts are
//            prefix-summed across the workgroup and each thread appends its
//            cell's grains (centre, r^2) at its offset, so the grains of cells
//            (i, j0..j1) are one contiguous LDS range;
//   phase 2  each thread owns (pixel, sample-slice) work items, walks the
//            window's columns and tests the contiguous grain ranges, stopping
//            at the first hit exactly like the reference's early exit;
//   phase 3  per-pixel hit counts are combined in LDS and stored once.
// If a tile's neighbourhood does not fit in LDS it is split in halves and each
// half processed in turn (down to a single pixel).
//
// Tile neighbourhood bounds are exact: every step from x to the window bound is
// a monotone rounding function, so evaluating it at the tile's extreme pixel and
// extreme offset bounds every point's window.
#include "fg_kernels.h"
#include "fg_rng.h"

namespace fg {

static __constant__ double c_zx[257] = {FG_ZIG_NORM_X_VALUES};
static __constant__ double c_zf[257] = {FG_ZIG_NORM_F_VALUES};

// floor(RN(a / delta)) as Rust's `as i32` (IEEE round-to-nearest division).
__device__ __forceinline__ int floor_div(float a, float delta)
{
    return sat_i32(floorf(__fdiv_rn(a, delta)));
}

// Workgroup exclusive scan of one u32 per thread; returns the exclusive prefix
// and the workgroup total.  Contains two barriers.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *s_wave,
                                                         uint32_t &total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t t = s_wave[w];
        before += (w < wid) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return before + x - v;
}

template <int KIND, bool STATS>
__global__ __launch_bounds__(kBlock) void pixelwise_tiles(PixelArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t s_cnt[kMaxTilePixels];
    __shared__ int4 s_stack[48];
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ int s_sp;

    const int tid = threadIdx.x;
    int tile = blockIdx.x;
    if (a.tile_list) tile = a.tile_list[tile];
    if (tile >= a.n_tiles) return;
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    const float rm = a.rm, delta = a.delta, iz = a.inv_zoom;
    const float4 orange = *a.off_range;

    for (int p = tid; p < kMaxTilePixels; p += kBlock) s_cnt[p] = 0;
    if (tid == 0) {
        s_stack[0] = make_int4(tx * a.tile_w, ty * a.tile_h, min((tx + 1) * a.tile_w, a.out_w),
                               min((ty + 1) * a.tile_h, a.out_h));
        s_sp = 1;
    }
    __syncthreads();

    unsigned long long st_points = 0, st_tests = 0, st_hits = 0;

    for (;;) {
        const int sp = s_sp;
        if (sp == 0) break;
        const int4 r = s_stack[sp - 1];
        __syncthreads();
        if (tid == 0) s_sp = sp - 1;

        // ---- neighbourhood of the rect (exact bounds, see header) ----
        const float px_lo = ((float)r.x + 0.5f) * iz, px_hi = ((float)(r.z - 1) + 0.5f) * iz;
        const float py_lo = ((float)r.y + 0.5f) * iz, py_hi = ((float)(r.w - 1) + 0.5f) * iz;
        const int I0 = floor_div((px_lo - orange.y) - rm, delta);
        const int I1 = floor_div((px_hi - orange.x) + rm, delta);
        const int J0 = floor_div((py_lo - orange.w) - rm, delta);
        const int J1 = floor_div((py_hi - orange.z) + rm, delta);
        const long long W = (long long)I1 - I0 + 1, H = (long long)J1 - J0 + 1;
        const long long C64 = W * H;
        const int gsize = KIND == 0 ? 8 : 12;
        const long long prefix_bytes = ((C64 + 1) * 4 + 15) & ~15LL;
        bool fits = W > 0 && H > 0 && prefix_bytes + 64LL * gsize <= a.lds_bytes;

        uint32_t *prefix = reinterpret_cast<uint32_t *>(lds);
        uint32_t cap_g = 0;
        float2 *gpos = nullptr;
        float *gr2 = nullptr;
        const int C = fits ? (int)C64 : 0;
        const int Hc = (int)(H > 0 ? H : 1);
        if (fits) {
            cap_g = (uint32_t)((a.lds_bytes - prefix_bytes) / gsize);
            gpos = reinterpret_cast<float2 *>(lds + prefix_bytes);
            gr2 = reinterpret_cast<float *>(lds + prefix_bytes + 8LL * cap_g);
        }

        // ---- phase 1: generate the neighbourhood's cells into LDS ----
        unsigned long long running = 0;
        for (int cb = 0; fits && cb < C; cb += kBlock) {
            const int c = cb + tid;
            uint32_t q = 0;
            Xoshiro rng;
            rng.s0 = rng.s1 = rng.s2 = rng.s3 = 0;
            float sx = 0.0f, sy = 0.0f;
            if (c < C) {
                const int ci = c / Hc, cj = c - ci * Hc;
                const int i = I0 + ci, j = J0 + cj;
                rng.seed(mix3(a.seed, kCellStream, i, j));
                sx = (float)i * delta;
                sy = (float)j * delta;
                // Plane::get_clamped(floor(sx), floor(sy)) (src/model.rs:60-67)
                const float fx = floorf(sx), fy = floorf(sy);
                const int ix = fx < 0.0f ? 0 : (fx > (float)(a.in_w - 1) ? a.in_w - 1 : (int)fx);
                const int iy = fy < 0.0f ? 0 : (fy > (float)(a.in_h - 1) ? a.in_h - 1 : (int)fy);
                const float lam = a.lambda[(size_t)iy * a.in_w + ix];
                if (lam > 0.0f) {
                    const float expected = lam * delta * delta;
                    if (expected > 0.0f) q = poisson_u32(rng, (double)expected);
                }
            }
            uint32_t total;
            const uint32_t excl = block_exclusive_scan(q, s_wave, total);
            if (running + total > cap_g) {  // uniform: running and total are workgroup-wide
                fits = false;
                break;
            }
            const uint32_t base = (uint32_t)running + excl;
            if (c < C) prefix[c] = base;
            for (uint32_t t = 0; t < q; ++t) {
                const float cx = sx + uniform_f32(rng, a.uscale);
                const float cy = sy + uniform_f32(rng, a.uscale);
                gpos[base + t] = make_float2(cx, cy);
                if (KIND == 1) {
                    float rad = lognormal_f32(rng, a.log_mu, a.log_sigma, c_zx, c_zf);
                    if (rad > rm) rad = rm;
                    gr2[base + t] = rad <= 0.0f ? -1.0f : rad * rad;
                }
            }
            running += total;
        }

        if (!fits) {
            // Split the rect in halves along its longer side; a single pixel that
            // does not fit is reported and left unwritten.
            if (tid == 0) {
                const int w = r.z - r.x, h = r.w - r.y;
                int top = s_sp;
                if (w * h <= 1 || top + 2 > 48) {
                    atomicOr(a.status, kStatusNeighbourhood);
                } else if (w >= h) {
                    const int xm = r.x + w / 2;
                    s_stack[top++] = make_int4(xm, r.y, r.z, r.w);
                    s_stack[top++] = make_int4(r.x, r.y, xm, r.w);
                } else {
                    const int ym = r.y + h / 2;
                    s_stack[top++] = make_int4(r.x, ym, r.z, r.w);
                    s_stack[top++] = make_int4(r.x, r.y, r.z, ym);
                }
                s_sp = top;
            }
            __syncthreads();
            continue;
        }
        if (tid == 0) prefix[C] = (uint32_t)running;
        __syncthreads();

        // ---- phase 2: coverage tests ----
        const int rw = r.z - r.x, rh = r.w - r.y, P = rw * rh;
        const int S = P >= kBlock ? 1 : kBlock / P;
        const int items = P * S;
        const float r2c = a.r2_const;
        for (int item = tid; item < items; item += kBlock) {
            const int p = item % P, sl = item / P;
            const int x = r.x + p % rw, y = r.y + p / rw;
            const float pxx = ((float)x + 0.5f) * iz;
            const float pyy = ((float)y + 0.5f) * iz;
            uint32_t cnt = 0;
            for (int k = sl; k < a.n_samples; k += S) {
                const float2 o = a.offs[k];
                const float xg = pxx - o.x, yg = pyy - o.y;
                const int i0 = max(floor_div(xg - rm, delta), I0);
                const int i1 = min(floor_div(xg + rm, delta), I1);
                const int j0 = max(floor_div(yg - rm, delta), J0);
                const int j1 = min(floor_div(yg + rm, delta), J1);
                bool hit = false;
                // j0 <= j1 always holds for finite offsets; the guard keeps the
                // prefix reads in range for any input.
                for (int i = i0; i <= i1 && j0 <= j1 && !hit; ++i) {
                    const int col = (i - I0) * Hc - J0;
                    const uint32_t gs = prefix[col + j0], ge = prefix[col + j1 + 1];
                    for (uint32_t g = gs; g < ge; ++g) {
                        const float2 gp = gpos[g];
                        const float r2 = KIND == 0 ? r2c : gr2[g];
                        const float dx = xg - gp.x, dy = yg - gp.y;
                        if (STATS && r2 >= 0.0f) ++st_tests;
                        if (dx * dx + dy * dy <= r2) {
                            hit = true;
                            break;
                        }
                    }
                }
                cnt += hit ? 1u : 0u;
                if (STATS) ++st_points;
            }
            if (STATS) st_hits += cnt;
            atomicAdd(&s_cnt[p], cnt);
        }
        __syncthreads();
        // ---- phase 3: store ----
        for (int p = tid; p < P; p += kBlock) {
            const int x = r.x + p % rw, y = r.y + p / rw;
            a.out[(size_t)y * a.out_w + x] = (float)s_cnt[p] * a.inv_samples;
            s_cnt[p] = 0;
        }
        __syncthreads();
    }
    if (STATS) {
        atomicAdd(&a.stats[0], st_points);
        atomicAdd(&a.stats[1], st_tests);
        atomicAdd(&a.stats[2], st_hits);
    }
}

__global__ __launch_bounds__(kBlock) void offsets_range_kernel(const float2 *offs, int n, float4 *range)
{
    __shared__ float4 s_r[kBlock];
    float4 r = make_float4(INFINITY, -INFINITY, INFINITY, -INFINITY);
    for (int k = threadIdx.x; k < n; k += kBlock) {
        const float2 o = offs[k];
        r.x = fminf(r.x, o.x);
        r.y = fmaxf(r.y, o.x);
        r.z = fminf(r.z, o.y);
        r.w = fmaxf(r.w, o.y);
    }
    s_r[threadIdx.x] = r;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const float4 b = s_r[threadIdx.x + w];
            float4 c = s_r[threadIdx.x];
            c.x = fminf(c.x, b.x);
            c.y = fmaxf(c.y, b.y);
            c.z = fminf(c.z, b.z);
            c.w = fmaxf(c.w, b.w);
            s_r[threadIdx.x] = c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *range = s_r[0];
}

hipError_t launch_offsets_range(const float2 *offs, int n, float4 *range, hipStream_t s)
{
    hipLaunchKernelGGL(offsets_range_kernel, dim3(1), dim3(kBlock), 0, s, offs, n, range);
    return hipGetLastError();
}

hipError_t launch_pixelwise(const PixelArgs &a, int dist_kind, bool with_stats, hipStream_t s)
{
    const dim3 grid(a.n_tiles), block(kBlock);
    const size_t lds = (size_t)a.lds_bytes;
    if (dist_kind == 0) {
        if (with_stats)
            hipLaunchKernelGGL((pixelwise_tiles<0, true>), grid, block, lds, s, a);
        else
            hipLaunchKernelGGL((pixelwise_tiles<0, false>), grid, block, lds, s, a);
    } else {
        if (with_stats)
            hipLaunchKernelGGL((pixelwise_tiles<1, true>), grid, block, lds, s, a);
        else
            hipLaunchKernelGGL((pixelwise_tiles<1, false>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}

}  // namespace fg
