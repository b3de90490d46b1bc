// fg_pixelwise.hip -- gfx950 pixel-wise Boolean-model integrator.
//
// Computes exactly what render_pixelwise / evaluate_indicator compute
// (src/pixelwise.rs:11-106): for every output pixel and Monte-Carlo offset k,
// the point (xg, yg) = ((x+0.5)/zoom - off_k.x, (y+0.5)/zoom - off_k.y) is
// covered iff some grain of some cell (i, j) in the reference's window
// [floor((xg-rm)/delta), floor((xg+rm)/delta)] x [...] contains it
// (dx*dx + dy*dy <= r*r in f32); the pixel value is hits * (1/N).
//
// The reference regenerates a cell's grains at every (pixel, sample, cell)
// visit.  A cell's grains depend only on (seed, i, j, lambda(cell)), so each
// workgroup here owns an output tile and generates every cell of the tile's
// neighbourhood once, in LDS.  The neighbourhood bound is exact: each step from
// a pixel index to its window bound is a monotone rounding function, so the
// tile's extreme pixel and extreme offset bound every point's window.
//
// Two formulations, bit-identical by construction:
//   splat   (default) grains are tested per sample: for grain g and offset k the
//           output pixels whose sample point can lie in g's disk form a box of at
//           most a few pixels; candidate (g, k, box) entries go to a per-wave LDS
//           queue and are drained 64 at a time through the exact reference test
//           (disk test plus window membership of g's cell), setting bit k of
//           the pixel in an LDS bitset; a popcount epilogue stores the tile.
//   gather  each (pixel, sample) walks its window's cells in the reference's
//           order and stops at the first hit; kept as the instrumented twin that
//           counts the reference's disk tests T (the FP32 FLOP count 16 + 6T).
#include <cmath>
#include <cstring>

#include "fg_device.h"

namespace fg {

// ---------------------------------------------------------------------------
// gather formulation
// ---------------------------------------------------------------------------
template <int KIND, bool STATS>
__global__ __launch_bounds__(kBlock) void pixelwise_gather(PixelArgs a)
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
    const float rm = a.rm, delta = a.delta, rd = a.rdelta, iz = a.inv_zoom;
    const int fast = a.div_fast;
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

        // ---- neighbourhood of the rect (exact bounds) ----
        const float px_lo = ((float)r.x + 0.5f) * iz, px_hi = ((float)(r.z - 1) + 0.5f) * iz;
        const float py_lo = ((float)r.y + 0.5f) * iz, py_hi = ((float)(r.w - 1) + 0.5f) * iz;
        const int I0 = floor_div((px_lo - orange.y) - rm, delta, rd, fast);
        const int I1 = floor_div((px_hi - orange.x) + rm, delta, rd, fast);
        const int J0 = floor_div((py_lo - orange.w) - rm, delta, rd, fast);
        const int J1 = floor_div((py_hi - orange.z) + rm, delta, rd, fast);
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

        // ---- phase 1: cells in column-major order (i outer, j inner) ----
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
                sx = (float)i * delta;
                sy = (float)j * delta;
                // Plane::get_clamped(floor(sx), floor(sy)) (src/model.rs:60-67)
                const float fx = floorf(sx), fy = floorf(sy);
                const int ix = fx < 0.0f ? 0 : (fx > (float)(a.in_w - 1) ? a.in_w - 1 : (int)fx);
                const int iy = fy < 0.0f ? 0 : (fy > (float)(a.in_h - 1) ? a.in_h - 1 : (int)fy);
                const double e = cell_threshold(a.lambda[(size_t)iy * a.in_w + ix], delta);
                q = cell_count(mix3(a.seed, kCellStream, i, j), e, rng);
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
                    float rad = lognormal_f32(rng, a.log_mu, a.log_sigma, d_zig_x, d_zig_f);
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

        // ---- phase 2: coverage tests in the reference's order ----
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
                const int i0 = max(floor_div(xg - rm, delta, rd, fast), I0);
                const int i1 = min(floor_div(xg + rm, delta, rd, fast), I1);
                const int j0 = max(floor_div(yg - rm, delta, rd, fast), J0);
                const int j1 = min(floor_div(yg + rm, delta, rd, fast), J1);
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

// ---------------------------------------------------------------------------
// splat formulation
// ---------------------------------------------------------------------------
template <int KIND>
__global__ __launch_bounds__(kBlock) void pixelwise_splat(PixelArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t s_queue[kBlock / 64][kQueueCap];
    __shared__ uint32_t s_wave[kBlock / 64];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int tile = blockIdx.x;
    if (a.tile_list) tile = a.tile_list[tile];
    if (tile >= a.n_tiles) return;
    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    SplatTile t;
    t.X0 = txi * a.tile_w;
    t.Y0 = tyi * a.tile_h;
    t.X1 = min(t.X0 + a.tile_w, a.out_w);
    t.Y1 = min(t.Y0 + a.tile_h, a.out_h);
    t.TW = a.tile_w;
    t.fx0 = (float)t.X0;
    t.fx1 = (float)(t.X1 - 1);
    t.fy0 = (float)t.Y0;
    t.fy1 = (float)(t.Y1 - 1);
    const float rm = a.rm, delta = a.delta, rd = a.rdelta, iz = a.inv_zoom, z = a.zoom;
    const int fast = a.div_fast;
    const int N = a.n_samples, words = a.words;
    const float4 orange = *a.off_range;

    // ---- neighbourhood of the tile (exact bounds) ----
    const float px_lo = (t.fx0 + 0.5f) * iz, px_hi = (t.fx1 + 0.5f) * iz;
    const float py_lo = (t.fy0 + 0.5f) * iz, py_hi = (t.fy1 + 0.5f) * iz;
    const int I0 = floor_div((px_lo - orange.y) - rm, delta, rd, fast);
    const int I1 = floor_div((px_hi - orange.x) + rm, delta, rd, fast);
    const int J0 = floor_div((py_lo - orange.w) - rm, delta, rd, fast);
    const int J1 = floor_div((py_hi - orange.z) + rm, delta, rd, fast);
    t.I0 = I0;
    t.J0 = J0;
    const long long W = (long long)I1 - I0 + 1, H = (long long)J1 - J0 + 1;
    if (W <= 0 || H <= 0 || W > 65535 || H > 65535 || W * H > 0x7fffffffLL) {
        if (tid == 0) atomicOr(a.status, kStatusRegion);
        return;
    }
    const int C = (int)(W * H);

    // ---- LDS layout ----
    size_t off = 0;
    uint32_t *bits = reinterpret_cast<uint32_t *>(lds);
    const int tile_words = a.tile_w * a.tile_h * words;
    off += ((size_t)tile_words * 4 + 15) & ~(size_t)15;
    const float2 *offs = a.offs;
    if (a.offs_in_lds) {
        float2 *ol = reinterpret_cast<float2 *>(lds + off);
        for (int k = tid; k < N; k += kBlock) ol[k] = a.offs[k];
        offs = ol;
        off += ((size_t)N * 8 + 15) & ~(size_t)15;
    }
    // lambda -> Poisson threshold of the region's input pixels
    const int PX0 = (int)fminf(fmaxf(floorf((float)I0 * delta), 0.0f), (float)(a.in_w - 1));
    const int PX1 = (int)fminf(fmaxf(floorf((float)I1 * delta), 0.0f), (float)(a.in_w - 1));
    const int PY0 = (int)fminf(fmaxf(floorf((float)J0 * delta), 0.0f), (float)(a.in_h - 1));
    const int PY1 = (int)fminf(fmaxf(floorf((float)J1 * delta), 0.0f), (float)(a.in_h - 1));
    const int ETW = PX1 - PX0 + 1, ETH = PY1 - PY0 + 1;
    const size_t et_bytes = (((size_t)ETW * ETH * 8) + 15) & ~(size_t)15;
    const int gsize = 32;
    const bool use_et = off + et_bytes + 256 * gsize <= (size_t)a.lds_bytes;
    double *etab = reinterpret_cast<double *>(lds + off);
    if (use_et) {
        for (int p = tid; p < ETW * ETH; p += kBlock) {
            const int px = PX0 + p % ETW, py = PY0 + p / ETW;
            etab[p] = cell_threshold(a.lambda[(size_t)py * a.in_w + px], delta);
        }
        off += et_bytes;
    }
    const int gcap = (int)min((long long)((a.lds_bytes - (long long)off) / gsize), 65535LL);
    if (gcap < kBlock) {  // the host sizes LDS so that one round of grains always fits
        if (tid == 0) atomicOr(a.status, kStatusRegion);
        return;
    }
    float4 *gA = reinterpret_cast<float4 *>(lds + off);
    float4 *gB = gA + gcap;
    for (int w = tid; w < tile_words; w += kBlock) bits[w] = 0u;

    // sample mapping: lanes over samples k, groups over grains
    const int KL = N >= kBlock ? kBlock : N;
    const int GS = kBlock / KL;
    const int kl = tid % KL, gsub = tid / KL;
    const bool tester = tid < KL * GS;
    const int KM = (N + KL - 1) / KL;  // samples per thread
    float oxz0 = 0.0f, oyz0 = 0.0f;
    if (KM == 1 && tester) {
        const float2 o = a.offs[kl];
        oxz0 = o.x * z;
        oyz0 = o.y * z;
    }
    // candidate box slack: covers the rounding of xg, dx and of the box itself
    const float slack = 4e-3f + 2e-6f * (float)max(a.out_w, a.out_h);
    uint32_t *queue = s_queue[wid];
    int qcount = 0;  // wave-uniform
    const bool one_k = KM == 1;
    const bool kvalid = tester && kl < N;
    const uint32_t klane = (uint32_t)kl << 16;
    __syncthreads();

    auto push = [&](bool cand, uint32_t entry) {
        const unsigned long long mask = __ballot(cand);
        if (cand) queue[qcount + __popcll(mask & ((1ull << lane) - 1ull))] = entry;
        qcount += __popcll(mask);
        if (qcount >= 64) {
            splat_drain(a, t, queue, 64, qcount, gA, gB, offs, bits);
            qcount -= 64;
        }
    };

    auto flush = [&](int fill) {
        __syncthreads();
        const int TT = (fill + GS - 1) / GS;
        for (int it = 0; it < TT; ++it) {
            const int g = gsub + it * GS;
            const bool gv = tester && g < fill;
            const float4 B = gB[gv ? g : 0];
            if (one_k) {
                const SplatBox b = splat_box(B, oxz0, oyz0);
                push(gv && kvalid && splat_box_hits(b, t), (uint32_t)g | klane);
            } else {
                for (int m = 0; m < KM; ++m) {
                    const int k = kl + m * KL;
                    const float2 o = offs[k < N ? k : 0];
                    const SplatBox b = splat_box(B, o.x * z, o.y * z);
                    push(gv && k < N && splat_box_hits(b, t), (uint32_t)g | ((uint32_t)k << 16));
                }
            }
        }
        if (qcount > 0) {
            splat_drain(a, t, queue, qcount, qcount, gA, gB, offs, bits);
            qcount = 0;
        }
        __syncthreads();
    };

    // ---- generation: cells in row-major order, grains batched through LDS ----
    const int Wi = (int)W;
    const float zs = z;
    int fill = 0;  // uniform
    // grains one thread may add per round, so that one round always fits the
    // buffer (gcap >= kBlock is guaranteed by the host's LDS sizing)
    const uint32_t kRound = (uint32_t)max(1, min(8, gcap / kBlock));
    for (int cb = 0; cb < C; cb += kBlock) {
        const int c = cb + tid;
        uint32_t rem = 0;
        Xoshiro rng;
        rng.s0 = rng.s1 = rng.s2 = rng.s3 = 0;
        float sx = 0.0f, sy = 0.0f;
        uint32_t cellpk = 0;
        if (c < C) {
            const int cj = c / Wi, ci = c - cj * Wi;
            const int i = I0 + ci, j = J0 + cj;
            sx = (float)i * delta;
            sy = (float)j * delta;
            const float fx = floorf(sx), fy = floorf(sy);
            const int ix = fx < 0.0f ? 0 : (fx > (float)(a.in_w - 1) ? a.in_w - 1 : (int)fx);
            const int iy = fy < 0.0f ? 0 : (fy > (float)(a.in_h - 1) ? a.in_h - 1 : (int)fy);
            const double e = use_et ? etab[(iy - PY0) * ETW + (ix - PX0)]
                                    : cell_threshold(a.lambda[(size_t)iy * a.in_w + ix], delta);
            rem = cell_count(mix3(a.seed, kCellStream, i, j), e, rng);
            cellpk = (uint32_t)ci | ((uint32_t)cj << 16);
        }
        while (__syncthreads_or(rem > 0u)) {
            const uint32_t n = rem < kRound ? rem : kRound;
            uint32_t total;
            const uint32_t excl = block_exclusive_scan(n, s_wave, total);
            if (fill + (int)total > gcap) {
                flush(fill);
                fill = 0;
            }
            const int base = fill + (int)excl;
            for (uint32_t q = 0; q < n; ++q) {
                const float cx = sx + uniform_f32(rng, a.uscale);
                const float cy = sy + uniform_f32(rng, a.uscale);
                float rad = a.r_const;
                if (KIND == 1) {
                    rad = lognormal_f32(rng, a.log_mu, a.log_sigma, d_zig_x, d_zig_f);
                    if (rad > rm) rad = rm;
                }
                const float r2 = rad <= 0.0f ? -1.0f : rad * rad;
                const float rz = fmaxf(rad, 0.0f) * zs + slack;
                gA[base + q] = make_float4(cx, cy, r2, __uint_as_float(cellpk));
                gB[base + q] = make_float4(cx * zs - 0.5f, cy * zs - 0.5f, rz, 0.0f);
            }
            fill += (int)total;
            rem -= n;
        }
    }
    flush(fill);

    // ---- popcount epilogue: one store per output pixel ----
    const int tw = t.X1 - t.X0, th = t.Y1 - t.Y0;
    for (int p = tid; p < tw * th; p += kBlock) {
        const int ox = p % tw, oy = p / tw;
        const uint32_t *pw = bits + ((size_t)oy * a.tile_w + ox) * words;
        uint32_t count = 0;
        for (int w = 0; w < words; ++w) count += __popc(pw[w]);
        a.out[(size_t)(t.Y0 + oy) * a.out_w + (t.X0 + ox)] = (float)count * a.inv_samples;
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

hipError_t launch_pixelwise(const PixelArgs &a, int dist_kind, int variant, bool with_stats, hipStream_t s)
{
    const dim3 grid(a.n_tiles), block(kBlock);
    const size_t lds = (size_t)a.lds_bytes;
    if (with_stats || variant == 1) {
        if (dist_kind == 0) {
            if (with_stats)
                hipLaunchKernelGGL((pixelwise_gather<0, true>), grid, block, lds, s, a);
            else
                hipLaunchKernelGGL((pixelwise_gather<0, false>), grid, block, lds, s, a);
        } else {
            if (with_stats)
                hipLaunchKernelGGL((pixelwise_gather<1, true>), grid, block, lds, s, a);
            else
                hipLaunchKernelGGL((pixelwise_gather<1, false>), grid, block, lds, s, a);
        }
    } else {
        if (dist_kind == 0)
            hipLaunchKernelGGL((pixelwise_splat<0>), grid, block, lds, s, a);
        else
            hipLaunchKernelGGL((pixelwise_splat<1>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}

// q0 = a*rd; r = fma(-q0, d, a); q = fma(r, rd, q0) against a/d over a in [1, 2).
bool reciprocal_division_exact(float d, float rd)
{
    if (!(d > 0.0f) || !std::isfinite(d) || !(rd > 0.0f) || !std::isfinite(rd)) return false;
    uint32_t bits = 0x3F800000u;
    for (uint32_t m = 0; m < (1u << 23); ++m, ++bits) {
        float a;
        std::memcpy(&a, &bits, 4);
        const volatile float q0 = a * rd;
        const float r = std::fma(-(float)q0, d, a);
        const float q = std::fma(r, rd, (float)q0);
        const volatile float exact = a / d;
        if (q != exact) return false;
    }
    return true;
}

}  // namespace fg
