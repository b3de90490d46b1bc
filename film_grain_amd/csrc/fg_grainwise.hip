// fg_grainwise.hip -- gfx950 grain-wise Boolean-model integrator.
//
// Computes exactly what render_grainwise computes (src/grainwise.rs:12-142):
// every input pixel (x, y) with lambda > 0 seeds pixel_rng(seed, x, y), draws
// q ~ Poisson(lambda) grains (centre (x + U, y + U), radius clamped to rm),
// and for every Monte-Carlo offset k each grain sets bit k of the output pixels
// whose centres (o + 0.5) lie in the disk of radius r*zoom around
// (cx*zoom + off_k.x, cy*zoom + off_k.y), using the reference's bounds()
// rounding.  The output is popcount(bits) * (1/N).
//
// Each workgroup owns an output tile and keeps the tile's N-bit sample sets in
// LDS, so the per-hit atomics are LDS atomics and the popcount epilogue writes
// each output pixel once.  It regenerates the grains of every input pixel whose
// grains can reach the tile (tile / zoom plus a halo of rm*zoom + max|offset|).
// Grain generation and grain testing alternate in rounds: each thread draws up
// to grains_per_thread grains of its input pixel into its LDS slots (keeping
// its Xoshiro state in registers between rounds), then all threads test all
// slotted grains with lanes over samples k, so the 64 lanes of a wave test one
// grain against 64 different offsets and the grain is a uniform (broadcast)
// LDS read.
#include <algorithm>

#include "fg_device.h"
#include "fg_pixelwise_hbm.h"

namespace fg {

// src/grainwise.rs:126-142, with the image limit; returns false for None.
__device__ __forceinline__ bool ref_bounds(float center, float radius, int limit, int &lo, int &hi)
{
    int mn = sat_i32(ceilf((center - radius) - 0.5f));
    int mx = sat_i32(floorf((center + radius) - 0.5f));
    if (mx < mn) return false;
    if (limit <= 0) return false;
    const int last = limit - 1;
    if (mn > last || mx < 0) return false;
    mn = min(max(mn, 0), last);
    mx = min(max(mx, 0), last);
    if (mn > mx) return false;
    lo = mn;
    hi = mx;
    return true;
}

template <int KIND, bool STATS>
__global__ __launch_bounds__(kBlock) void grainwise_tiles(GrainArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t s_slot_n[kBlock];

    const int tid = threadIdx.x;
    const int tile = blockIdx.x;
    if (tile >= a.n_tiles) return;
    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    const int X0 = txi * a.tile_w, Y0 = tyi * a.tile_h;
    const int X1 = min(X0 + a.tile_w, a.out_w), Y1 = min(Y0 + a.tile_h, a.out_h);
    const int words = a.words;
    const int gpt = a.grains_per_thread;
    const int tile_px = a.tile_w * a.tile_h;

    uint32_t *bits = reinterpret_cast<uint32_t *>(lds);
    const size_t bits_bytes = (((size_t)tile_px * words * 4) + 15) & ~(size_t)15;
    float4 *slots = reinterpret_cast<float4 *>(lds + bits_bytes);

    for (int w = tid; w < tile_px * words; w += kBlock) bits[w] = 0u;

    // Input pixels whose grains can reach [X0, X1) x [Y0, Y1): centre in
    // [x, x+1], target t = cx*zoom + off, covered outputs within t +- R - 0.5.
    // The margins are generous; extra pixels only produce misses.
    const float z = a.zoom, R = a.r_out_max;
    const float4 orange = *a.off_range;
    const int xi0 = max(0, (int)floorf(((float)X0 + 0.5f - R - orange.y) / z) - 2);
    const int xi1 = min(a.in_w - 1, (int)ceilf(((float)X1 - 0.5f + R - orange.x) / z) + 2);
    const int yi0 = max(0, (int)floorf(((float)Y0 + 0.5f - R - orange.w) / z) - 2);
    const int yi1 = min(a.in_h - 1, (int)ceilf(((float)Y1 - 0.5f + R - orange.z) / z) + 2);
    const int rwi = xi1 - xi0 + 1;
    const int npx = (xi1 >= xi0 && yi1 >= yi0) ? rwi * (yi1 - yi0 + 1) : 0;

    // Test-phase mapping: lanes over samples, groups over grains.
    const int KL = a.n_samples >= kBlock ? kBlock : a.n_samples;
    const int GS = kBlock / KL;
    const int kl = tid % KL, gs = tid / KL;
    const bool tester = tid < KL * GS;

    const float uscale1 = 1.0f;  // Uniform::new(0.0f32, 1.0) keeps scale 1
    // instrumented twin: disk tests and hits of the tile's own pixels (each
    // output pixel lies in one tile, so the sums over tiles are the
    // reference's); grains and pairs are counted once per grain by
    // grain_count_kernel, since tiles regenerate their halo's grains
    unsigned long long st_tests = 0, st_hits = 0;
    __syncthreads();

    for (int pb = 0; pb < npx; pb += kBlock) {
        const int pix = pb + tid;
        Xoshiro rng;
        rng.s0 = rng.s1 = rng.s2 = rng.s3 = 0;
        uint32_t q = 0, done = 0;
        float fx = 0.0f, fy = 0.0f;
        if (pix < npx) {
            const int x = xi0 + pix % rwi, y = yi0 + pix / rwi;
            const float lam = a.lambda[(size_t)y * a.in_w + x];
            if (lambda_invalid(lam)) atomicOr(a.status, kStatusLambda);
            if (lam > 0.0f && lam <= kMaxLambda) {
                rng.seed(mix3(a.seed, kPixelStream, x, y));
                q = poisson_u32(rng, (double)lam);
            }
            fx = (float)x;
            fy = (float)y;
        }
        while (__syncthreads_or(done < q)) {
            int n = 0;
            while (n < gpt && done < q) {
                const float cx = fx + uniform_f32(rng, uscale1);
                const float cy = fy + uniform_f32(rng, uscale1);
                float rad = KIND == 0 ? a.r_const : lognormal_f32(rng, a.log_mu, a.log_sigma, d_zig_x, d_zig_f);
                ++done;
                if (rad > a.rm) rad = a.rm;
                if (rad <= 0.0f) continue;
                const float rout = rad * z;
                if (rout <= 0.0f) continue;
                slots[tid * gpt + n] = make_float4(cx * z, cy * z, rout, rout * rout);
                ++n;
            }
            s_slot_n[tid] = (uint32_t)n;
            __syncthreads();
            if (tester) {
                for (int b = 0; b < kBlock; ++b) {
                    const int nb = (int)s_slot_n[b];
                    for (int g = gs; g < nb; g += GS) {
                        const float4 G = slots[b * gpt + g];
                        for (int k = kl; k < a.n_samples; k += KL) {
                            const float2 o = a.offs[k];
                            const float txp = G.x + o.x, typ = G.y + o.y;
                            int xmn, xmx, ymn, ymx;
                            if (!ref_bounds(txp, G.z, a.out_w, xmn, xmx)) continue;
                            if (!ref_bounds(typ, G.z, a.out_h, ymn, ymx)) continue;
                            xmn = max(xmn, X0);
                            xmx = min(xmx, X1 - 1);
                            ymn = max(ymn, Y0);
                            ymx = min(ymx, Y1 - 1);
                            const uint32_t mask = 1u << (k & 31);
                            const int wk = k >> 5;
                            for (int oy = ymn; oy <= ymx; ++oy) {
                                const float dy = ((float)oy + 0.5f) - typ;
                                const float dy2 = dy * dy;
                                if (dy2 > G.w) continue;
                                for (int ox = xmn; ox <= xmx; ++ox) {
                                    const float dx = ((float)ox + 0.5f) - txp;
                                    if (STATS) ++st_tests;
                                    if (dx * dx + dy2 <= G.w) {
                                        atomicOr(&bits[((oy - Y0) * a.tile_w + (ox - X0)) * words + wk], mask);
                                        if (STATS) ++st_hits;
                                    }
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    }

    // popcount epilogue: one store per output pixel of the tile
    const int tw = X1 - X0, th = Y1 - Y0;
    for (int p = tid; p < tw * th; p += kBlock) {
        const int ox = p % tw, oy = p / tw;
        const uint32_t *pw = bits + ((size_t)oy * a.tile_w + ox) * words;
        uint32_t count = 0;
        for (int w = 0; w < words; ++w) count += __popc(pw[w]);
        a.out[(size_t)(Y0 + oy) * a.out_w + (X0 + ox)] = (float)count * a.inv_samples;
    }
    if (STATS) {
        atomicAdd(&a.stats[2], st_tests);
        atomicAdd(&a.stats[3], st_hits);
    }
}

// ---------------------------------------------------------------------------
// Sorted-sample formulation (same result, far fewer (grain, sample) pairs)
// ---------------------------------------------------------------------------
// For grain centre t = cx*zoom and sample k, bounds() has a column iff some
// integer n lies in [(t + X_k - R) - 0.5, (t + X_k + R) - 0.5], i.e. iff
// frac(X_k) lies in the cyclic interval [A - 2R, A] with A = frac(R - (t -
// 0.5)).  With the samples ordered by frac(X_k) (SampleTable, built from the
// output-pixel offsets) those samples are one cyclic run, found through the
// bucket index; the interval is widened by the table's eps, so the run holds
// every such sample.  For a constant radius with a narrow window, per-cell
// candidate lists over both fractions (cell_lists_kernel) replace the run.
// Each lane owns one input pixel of the tile's halo region: it draws that
// pixel's grains in the reference's order and, per grain, walks its list or
// run.  A (grain, sample) pair whose bounds() box, clipped to the tile, is one
// pixel takes the reference's disk test in the lane; larger boxes go to the
// wave's queue, one entry per pixel of the box inside the tile, drained
// through the same test.  Hits go to the tile's LDS bitset.
// Poisson pre-pass (src/grainwise.rs:37-45): per input pixel with lambda > 0,
// pixel_rng and the grain count; the count and the stream's state after the
// count draws are stored so that the splat kernel resumes each pixel's stream
// where the reference does (keeps the large Poisson sampler out of the splat
// kernel's registers).  Pixels with lambda <= 0 get q = 0.
__global__ __launch_bounds__(kBlock) void pixel_state_kernel(GrainArgs a)
{
    const long long n = (long long)a.in_w * a.in_h;
    for (long long p = (long long)blockIdx.x * kBlock + threadIdx.x; p < n; p += (long long)gridDim.x * kBlock) {
        const int y = (int)(p / a.in_w), x = (int)(p - (long long)y * a.in_w);
        const float lam = a.lambda[p];
        Xoshiro rng;
        rng.s0 = rng.s1 = rng.s2 = rng.s3 = 0;
        uint32_t q = 0;
        // values outside the accepted range are skipped and flagged (once per wave)
        if (__ballot(lambda_invalid(lam)) != 0ull && (threadIdx.x & 63) == 0) atomicOr(a.status, kStatusLambda);
        if (lam > 0.0f && lam <= kMaxLambda) {
            rng.seed(mix3(a.seed, kPixelStream, x, y));
            q = poisson_u32(rng, (double)lam);
        }
        a.pix_q[p] = q;
        a.pix_s01[p] = make_ulonglong2(rng.s0, rng.s1);
        a.pix_s23[p] = make_ulonglong2(rng.s2, rng.s3);
    }
}

// Instrumented twin, per input pixel: the grains the reference draws with a
// positive output radius (src/grainwise.rs:46-66), each of which it pairs with
// every sample.  stats[0] += grains, stats[1] += grains * N.
template <int KIND>
__global__ __launch_bounds__(kBlock) void grain_count_kernel(GrainArgs a)
{
    const long long n = (long long)a.in_w * a.in_h;
    unsigned long long cnt = 0;
    for (long long p = (long long)blockIdx.x * kBlock + threadIdx.x; p < n; p += (long long)gridDim.x * kBlock) {
        const int y = (int)(p / a.in_w), x = (int)(p - (long long)y * a.in_w);
        const float lam = a.lambda[p];
        if (!(lam > 0.0f && lam <= kMaxLambda)) continue;
        Xoshiro rng;
        rng.seed(mix3(a.seed, kPixelStream, x, y));
        const uint32_t q = poisson_u32(rng, (double)lam);
        for (uint32_t g = 0; g < q; ++g) {
            (void)uniform_f32(rng, 1.0f);
            (void)uniform_f32(rng, 1.0f);
            float rad = KIND == 0 ? a.r_const : lognormal_f32(rng, a.log_mu, a.log_sigma, d_zig_x, d_zig_f);
            if (rad > a.rm) rad = a.rm;
            if (rad <= 0.0f || rad * a.zoom <= 0.0f) continue;
            ++cnt;
        }
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((threadIdx.x & 63) == 0 && cnt) {
        atomicAdd(&a.stats[0], cnt);
        atomicAdd(&a.stats[1], cnt * (unsigned long long)a.n_samples);
    }
}

hipError_t launch_grain_count(const GrainArgs &a, int dist_kind, hipStream_t s)
{
    const long long n = (long long)a.in_w * a.in_h;
    const unsigned blocks = (unsigned)std::max<long long>(1, std::min<long long>((n + kBlock - 1) / kBlock, 8192));
    if (dist_kind == 0)
        hipLaunchKernelGGL(grain_count_kernel<0>, dim3(blocks), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(grain_count_kernel<1>, dim3(blocks), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pixel_state(const GrainArgs &a, hipStream_t s)
{
    const long long n = (long long)a.in_w * a.in_h;
    const long long blocks = std::min<long long>((n + kBlock - 1) / kBlock, 1LL << 20);
    hipLaunchKernelGGL(pixel_state_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// Input pixels whose grains can reach output rect [X0, X1) x [Y0, Y1): centre
// in [x, x+1], target cx*zoom + offset, covered outputs within target +- R -
// 0.5 (generous margins; extra pixels only produce grains that miss).  ext
// holds the offsets' (min x, max x, min y, max y).
struct PixRegion {
    int x0, y0, w, n;
};
__device__ __forceinline__ PixRegion grain_tile_region(const GrainArgs &a, const float4 &ext, int X0, int Y0, int X1,
                                                       int Y1)
{
    const float z = a.zoom, Rmax = a.r_out_max;
    const int xi0 = max(0, (int)floorf(((float)X0 + 0.5f - Rmax - ext.y) / z) - 2);
    const int xi1 = min(a.in_w - 1, (int)ceilf(((float)X1 - 0.5f + Rmax - ext.x) / z) + 2);
    const int yi0 = max(0, (int)floorf(((float)Y0 + 0.5f - Rmax - ext.w) / z) - 2);
    const int yi1 = min(a.in_h - 1, (int)ceilf(((float)Y1 - 0.5f + Rmax - ext.z) / z) + 2);
    PixRegion r;
    r.x0 = xi0;
    r.y0 = yi0;
    r.w = xi1 - xi0 + 1;
    r.n = (xi1 >= xi0 && yi1 >= yi0) ? r.w * (yi1 - yi0 + 1) : 0;
    return r;
}

// Work of each grain-wise tile (the grains of its input region, from the
// Poisson pre-pass), for a heaviest-first launch order; one workgroup per tile.
__global__ __launch_bounds__(kBlock) void grain_tile_weights_kernel(GrainArgs a, SampleTable st, uint32_t *w,
                                                                    int *ids)
{
    __shared__ uint32_t s_sum[kBlock / 64];
    const int tile = blockIdx.x;
    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    const int X0 = txi * a.tile_w, Y0 = tyi * a.tile_h;
    const PixRegion r = grain_tile_region(a, *st.ext, X0, Y0, min(X0 + a.tile_w, a.out_w),
                                          min(Y0 + a.tile_h, a.out_h));
    uint32_t sum = 0;
    for (int p = threadIdx.x; p < r.n; p += kBlock) {
        const int y = r.y0 + p / r.w, x = r.x0 + p % r.w;
        sum += a.pix_q[(size_t)y * a.in_w + x];
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kBlock / 64; ++i) t += s_sum[i];
        w[tile] = t;
        ids[tile] = tile;
    }
}

hipError_t launch_grain_tile_weights(const GrainArgs &a, const SampleTable &st, uint32_t *w, int *ids,
                                     hipStream_t s)
{
    hipLaunchKernelGGL(grain_tile_weights_kernel, dim3((unsigned)a.n_tiles), dim3(kBlock), 0, s, a, st, w, ids);
    return hipGetLastError();
}

// Entries of a round's concatenated candidate lists that the per-wave owner
// map holds (256 B per wave keeps a 64x32 tile's workgroup at 40 KB of LDS,
// four per CU).
constexpr int kOwnerMap = 256;

// Entry pos of a candidate list: a 32-bit byte offset from the list's base,
// so that the load takes the base in SGPRs and the offset in one VGPR (no
// 64-bit address arithmetic in the walk's steps: 19.55 -> 19.28 ms at C3).
__device__ __forceinline__ int list_at(const uint16_t *list, uint32_t pos)
{
    return (int)*reinterpret_cast<const uint16_t *>(reinterpret_cast<const unsigned char *>(list) + (pos << 1));
}

template <int KIND, int BLOCK>
__global__ __launch_bounds__(BLOCK, 8) void grainwise_sorted(GrainArgs a, SampleTable st)
{
    constexpr int NW = BLOCK / 64;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t s_queue[NW][kQueueCap];
    // per wave, the owner lane of each entry of a round's concatenated lists
    __shared__ uint8_t s_owner[NW][kOwnerMap];
    // the region's next pixel that no lane has taken
    __shared__ int s_next;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int tile = blockIdx.x;
    if (a.tile_list) tile = a.tile_list[tile];
    if (tile >= a.n_tiles) return;
    if (tile_cancelled(a.cancel, a.status)) return;
    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    const int X0 = txi * a.tile_w, Y0 = tyi * a.tile_h;
    const int X1 = min(X0 + a.tile_w, a.out_w), Y1 = min(Y0 + a.tile_h, a.out_h);
    // per-tile floats held in SGPRs (this target has no scalar float ALU, so
    // they are computed in VGPRs; kept there all kernel long, they spilled)
    const float fx0 = uniform_f((float)X0), fx1 = uniform_f((float)(X1 - 1));
    const float fy0 = uniform_f((float)Y0), fy1 = uniform_f((float)(Y1 - 1));
    const int words = a.words, N = a.n_samples, Q = st.qbuckets;
    const int tile_words = a.tile_w * a.tile_h * words;

    // ---- LDS: bitset | xy[N] | k[N] | qidx[Q + 1] ----
    size_t off = 0;
    uint32_t *bits = reinterpret_cast<uint32_t *>(lds);
    off += ((size_t)tile_words * 4 + 15) & ~(size_t)15;
    float2 *sxy = reinterpret_cast<float2 *>(lds + off);
    off += ((size_t)N * 8 + 15) & ~(size_t)15;
    uint16_t *sk = reinterpret_cast<uint16_t *>(lds + off);
    off += ((size_t)N * 2 + 15) & ~(size_t)15;
    uint16_t *sq = reinterpret_cast<uint16_t *>(lds + off);
    off += ((size_t)(Q + 1) * 2 + 15) & ~(size_t)15;
    if (off > (size_t)a.lds_bytes) {
        if (tid == 0) atomicOr(a.status, kStatusRegion);
        return;
    }
    for (int w = tid; w < tile_words; w += BLOCK) bits[w] = 0u;
    for (int k = tid; k < N; k += BLOCK) {
        sxy[k] = st.xy[k];
        sk[k] = st.k[k];
    }
    for (int q = tid; q <= Q; q += BLOCK) sq[q] = st.qidx[q];
    if (tid == 0) s_next = 0;
    const float4 ext0 = *st.ext;
    const float4 ext = make_float4(uniform_f(ext0.x), uniform_f(ext0.y), uniform_f(ext0.z), uniform_f(ext0.w));

    const float z = a.zoom, Rmax = a.r_out_max;
    const PixRegion reg = grain_tile_region(a, ext, X0, Y0, X1, Y1);
    const int xi0 = reg.x0, yi0 = reg.y0, rwi = reg.w, npx = reg.n;
    __syncthreads();

    // widening of the fraction interval: every coordinate of a box is below
    // C in magnitude, so each of its f32 steps is off by at most C * 2^-24
    const float C = (float)max(a.out_w, a.out_h) + fmaxf(fmaxf(fabsf(ext.x), fabsf(ext.y)),
                                                         fmaxf(fabsf(ext.z), fabsf(ext.w))) + Rmax + 8.0f;
    const float eps = uniform_f(C * 9.5367431640625e-07f + 1.52587890625e-05f);  // C * 2^-20 + 2^-16
    const float Qf = uniform_f((float)Q);
    uint32_t *queue = s_queue[wid];
    int qcount = 0;
    // the lane's current grain: target origin and output radius
    float gbx = 0.0f, gby = 0.0f, gR = 0.0f;

    // Entries: lane (6 bits) | sorted position << 6 (11 bits) | tile column
    // << 17 (7 bits) | tile row << 24 (8 bits).  The reference's disk test
    // (src/grainwise.rs:82-100) for entries [0, n).
    auto drain = [&](int n) {
        uint32_t e = 0;
        if (lane < n) e = queue[lane];
        if (lane + 64 < qcount) queue[lane] = queue[lane + 64];
        const int l = (int)(e & 63u);
        const float bx = __shfl(gbx, l, 64), by = __shfl(gby, l, 64), R = __shfl(gR, l, 64);
        if (lane >= n) return;
        const int p = (int)((e >> 6) & 0x7FFu), tx = (int)((e >> 17) & 0x7Fu), ty = (int)(e >> 24);
        const float2 X = sxy[p];
        const int k = sk[p];
        const float txp = bx + X.x, typ = by + X.y;
        const float dy = ((float)(Y0 + ty) + 0.5f) - typ;
        const float dy2 = dy * dy;
        const float dx = ((float)(X0 + tx) + 0.5f) - txp;
        if (dx * dx + dy2 <= R * R) atomicOr(&bits[(ty * a.tile_w + tx) * words + (k >> 5)], 1u << (k & 31));
    };
    auto push = [&](bool cand, uint32_t entry) {
        const unsigned long long mask = __ballot(cand);
        const int before = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        if (cand) queue[qcount + before] = entry;
        qcount += __popcll(mask);
        if (qcount >= 64) {
            drain(64);
            qcount -= 64;
        }
    };

    // Each lane draws the grains of one region pixel at a time; a lane whose
    // pixel has no grain left takes the region's next pixel, so the wave's
    // lanes stay busy until the region runs out (pixel counts differ widely).
    // Grains reach the bitset in another order than the reference's, which
    // changes nothing: the bitset is an OR of the tests.
    Xoshiro rng;
    rng.s0 = rng.s1 = rng.s2 = rng.s3 = 0;
    uint32_t q = 0, done = 0;
    float fx = 0.0f, fy = 0.0f;
    bool more = true;  // uniform: pixels of the region are left
    for (;;) {
        while (more) {
            const bool need = done >= q;
            const unsigned long long m = __ballot(need);
            if (m == 0) break;
            const int cnt = __popcll(m);
            int base = 0;
            if (lane == 0) base = atomicAdd(&s_next, cnt);
            base = __shfl(base, 0, 64);
            more = base + cnt < npx;
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            const int pix = base + rank;
            if (need && pix < npx) {
                const int y = yi0 + pix / rwi, x = xi0 + pix % rwi;
                const size_t pi = (size_t)y * a.in_w + x;
                q = a.pix_q[pi];
                const ulonglong2 s01 = a.pix_s01[pi], s23 = a.pix_s23[pi];
                rng.s0 = s01.x;
                rng.s1 = s01.y;
                rng.s2 = s23.x;
                rng.s3 = s23.y;
                done = 0;
                fx = (float)x;
                fy = (float)y;
            }
        }
        if (!__any(done < q)) break;
        {
            // the lane's next grain (src/grainwise.rs:48-67)
            int p0 = 0, len = 0;
            float bx = 0.0f, by = 0.0f, R = 0.0f;
            if (done < q) {
                const float cx = fx + uniform_f32(rng, 1.0f);
                const float cy = fy + uniform_f32(rng, 1.0f);
                float rad = KIND == 0 ? a.r_const : lognormal_f32(rng, a.log_mu, a.log_sigma, d_zig_x, d_zig_f);
                ++done;
                if (rad > a.rm) rad = a.rm;
                R = rad * z;
                bx = cx * z;
                by = cy * z;
                // skip radii <= 0 and grains whose boxes miss the tile for
                // every sample (bounds() is monotone in the target); one mask,
                // not a chain of branches
                const bool keep = ((int)(rad > 0.0f) & (int)(R > 0.0f) &
                                   (int)(floorf(((bx + ext.y) + R) - 0.5f) >= fx0) &
                                   (int)(ceilf(((bx + ext.x) - R) - 0.5f) <= fx1) &
                                   (int)(floorf(((by + ext.w) + R) - 0.5f) >= fy0) &
                                   (int)(ceilf(((by + ext.z) - R) - 0.5f) <= fy1)) != 0;
                if (keep && a.cell_q > 0) {
                    // the candidate list of the cell of (A, B), the fractions
                    // of the grain's x and y targets (build_cell_lists)
                    const float ncx = R - (bx - 0.5f), ncy = R - (by - 0.5f);
                    const float A = ncx - floorf(ncx), B = ncy - floorf(ncy);
                    const float cq = (float)a.cell_q;
                    const int qa = min(max((int)(A * cq), 0), a.cell_q - 1);
                    const int qb = min(max((int)(B * cq), 0), a.cell_q - 1);
                    const int c = qb * a.cell_q + qa;
                    p0 = (int)a.cell_start[c];
                    len = (int)a.cell_start[c + 1] - p0;
                } else if (keep) {
                    const float nc = R - (bx - 0.5f);
                    const float A = nc - floorf(nc);
                    const float lo = (A - (R + R)) - eps, hi = A + eps;
                    if (!(lo * Qf > -2.0f * Qf) || !(hi * Qf < 2.0f * Qf) ||
                        (int)floorf(hi * Qf) + 1 - (int)floorf(lo * Qf) >= Q) {
                        len = N;
                    } else {
                        int lq = (int)floorf(lo * Qf), hq = (int)floorf(hi * Qf) + 1;
                        if (lq < 0) {
                            lq += Q;
                            hq += Q;
                        }
                        lq = min(max(lq, 0), Q - 1);
                        hq = min(max(hq, lq + 1), 2 * Q);
                        p0 = sq[lq];
                        const int end = hq <= Q ? (int)sq[hq] : N + (int)sq[hq - Q];
                        len = end - p0;
                    }
                }
            }
            gbx = bx;
            gby = by;
            gR = R;
            // One candidate per lane: sample p (sorted position) against the
            // grain of lane gl, whose target origin and output radius are
            // (ebx, eby) and eR; ok false: nothing to test.
            auto visit = [&](bool ok, int p, int gl, float ebx, float eby, float eR) {
                const float2 X = sxy[p];
                const float txp = ebx + X.x, typ = eby + X.y;
                // bounds() of src/grainwise.rs:126-142 before its clamp; the
                // tile lies inside the image, so meeting the tile is the clamp
                const float xl = ceilf((txp - eR) - 0.5f), xh = floorf((txp + eR) - 0.5f);
                const float yl = ceilf((typ - eR) - 0.5f), yh = floorf((typ + eR) - 0.5f);
                const bool cand = ok & (xl <= xh) & (yl <= yh) & (xl <= fx1) & (xh >= fx0) & (yl <= fy1) &
                                  (yh >= fy0);
                // a candidate box of one pixel lies in the tile (it meets
                // it); wider boxes queue their pixels inside the tile
                const bool multi = cand & ((xh > xl) | (yh > yl));
                if (!__any(multi)) {
                    // one pixel per candidate: the drain's test of it, run
                    // here in the lane (same f32 expression)
                    if (cand) {
                        const int k = sk[p];
                        const int tx = (int)(xl - fx0), ty = (int)(yl - fy0);
                        const float dx = (xl + 0.5f) - txp, dy = (yl + 0.5f) - typ;
                        if (dx * dx + dy * dy <= eR * eR)
                            atomicOr(&bits[(ty * a.tile_w + tx) * words + (k >> 5)], 1u << (k & 31));
                    }
                } else {
                    const uint32_t base = (uint32_t)gl | ((uint32_t)p << 6);
                    const float cxl = fmaxf(xl, fx0), cxh = fminf(xh, fx1);
                    const float cyl = fmaxf(yl, fy0), cyh = fminf(yh, fy1);
                    const int tx = (int)(cxl - fx0), ty = (int)(cyl - fy0);
                    const int bw = cand ? (int)(cxh - cxl) + 1 : 0, bh = cand ? (int)(cyh - cyl) + 1 : 0;
                    const int nb = bw * bh;
                    for (int j = 0; __any(j < nb); ++j) {
                        const bool c = j < nb;
                        const int jy = c ? j / bw : 0, jx = c ? j - jy * bw : 0;
                        push(c, base | ((uint32_t)(tx + jx) << 17) | ((uint32_t)(ty + jy) << 24));
                    }
                }
            };
            if (a.cell_q > 0) {
                // The lanes' candidate lists, concatenated in lane order, are
                // walked 64 entries a step (the per-lane walk ran for the
                // wave's longest list).  Entry e belongs to the first lane g
                // whose running total incl_g exceeds e, i.e. g = the number
                // of lanes with incl <= e (a binary search over the
                // nondecreasing incl), and is its list entry p0 + e - excl.
                const uint32_t incl = wave_inclusive_sum((uint32_t)len);
                const int total = __builtin_amdgcn_readlane((int)incl, 63);
                const int excl = (int)incl - len;
                const int dl = p0 - excl;
                // Rounds of at most kOwnerMap entries (nearly all: about two
                // per grain at C3) write each entry's owner lane into the
                // wave's map, so that a step reads it with one load; longer
                // rounds search the running totals.  The wave's LDS accesses
                // run in order, so no barrier is needed.
                const bool mapped = total <= kOwnerMap;  // uniform
                if (mapped) {
                    for (int j = 0; j < len; ++j) s_owner[wid][excl + j] = (uint8_t)lane;
                    for (int s0 = 0; s0 < total; s0 += 64) {
                        const int e = s0 + lane;
                        const int g = s_owner[wid][min(e, total - 1)];
                        const bool ok = e < total;
                        const int ga = g << 2;
                        const uint32_t pos = (uint32_t)(from_lane_i(dl, ga) + e);
                        const int p = ok ? list_at(a.cell_list, pos) : 0;
                        visit(ok, p, g, from_lane_f(bx, ga), from_lane_f(by, ga), from_lane_f(R, ga));
                    }
                } else {
                    for (int s0 = 0; s0 < total; s0 += 64) {
                        const int e = s0 + lane;
                        int g = 0;
#pragma unroll
                        for (int b = 32; b >= 1; b >>= 1)
                            g += from_lane_i((int)incl, (g + b - 1) << 2) <= e ? b : 0;
                        const bool ok = e < total;  // then g <= 63 and its list holds e
                        const int ga = g << 2;
                        const uint32_t pos = (uint32_t)(from_lane_i(dl, ga) + e);
                        const int p = ok ? list_at(a.cell_list, pos) : 0;
                        visit(ok, p, g, from_lane_f(bx, ga), from_lane_f(by, ga), from_lane_f(R, ga));
                    }
                }
            } else {
                for (int it = 0; __any(it < len); ++it) {
                    // the cyclic run position (lanes past their run read a
                    // valid sample and queue nothing)
                    const int pr = p0 + it;
                    visit(it < len, (int)min(min((uint32_t)pr, (uint32_t)(pr - N)), (uint32_t)(N - 1)), lane, bx, by,
                          R);
                }
            }
            // the queue refers to the lanes' current grains
            if (qcount > 0) {
                drain(qcount);
                qcount = 0;
            }
        }
    }
    __syncthreads();

    const int tw = X1 - X0, th = Y1 - Y0;
    for (int p = tid; p < tw * th; p += BLOCK) {
        const int ox = p % tw, oy = p / tw;
        const uint32_t *pw = bits + ((size_t)oy * a.tile_w + ox) * words;
        uint32_t count = 0;
        for (int w = 0; w < words; ++w) count += __popc(pw[w]);
        a.out[(size_t)(Y0 + oy) * a.out_w + (X0 + ox)] = (float)count * a.inv_samples;
    }
}

// Candidate lists for grains of one constant output radius R: cell (qa, qb)
// of a Q x Q grid over the fractions (A, B) of a grain's x and y targets
// lists, in sorted position order, every sample whose fractions lie in
// [qa/Q - 2R - eps, (qa+1)/Q + eps] and [qb/Q - 2R - eps, (qb+1)/Q + eps]
// (mod 1), each widened by 2^-12.  A grain in the cell needs at most the
// samples whose fraction lies in [A - 2R - eps, A + eps] (its box then has a
// column: the per-grain run's argument) and in [B - 2R - eps, B + eps] (a
// row), so the list is a superset of the candidates and the exact tests
// downstream decide.  eps is the kernel's: C * 2^-20 + 2^-16 with C =
// c_base + the offsets' largest magnitude.  One workgroup, one cell per
// thread and pass; start has Q * Q + 1 entries, list room for Q * Q * N.
__global__ __launch_bounds__(1024) void cell_lists_kernel(SampleTable st, int N, int Q, float R, float c_base,
                                                          uint32_t *start, uint16_t *list)
{
    __shared__ uint32_t s_count[1024];
    const float4 ext = *st.ext;
    const float C = c_base + fmaxf(fmaxf(fabsf(ext.x), fabsf(ext.y)), fmaxf(fabsf(ext.z), fabsf(ext.w)));
    const float eps = C * 9.5367431640625e-07f + 1.52587890625e-05f;
    const float m = 2.44140625e-4f;  // 2^-12
    const float w = ((R + R) + (eps + eps)) + (1.0f / (float)Q + (m + m));
    const int cells = Q * Q;
    auto member = [&](int c, int p) {
        const float2 X = st.xy[p];
        const float fx = X.x - floorf(X.x), fy = X.y - floorf(X.y);
        const float lx = ((float)(c % Q) / (float)Q - (R + R)) - (eps + m);
        const float ly = ((float)(c / Q) / (float)Q - (R + R)) - (eps + m);
        float dx = fx - lx, dy = fy - ly;
        dx -= floorf(dx);
        dy -= floorf(dy);
        return dx <= w && dy <= w;
    };
    uint32_t base = 0;
    for (int c0 = 0; c0 < cells; c0 += 1024) {
        const int c = c0 + (int)threadIdx.x;
        uint32_t n = 0;
        if (c < cells)
            for (int p = 0; p < N; ++p) n += member(c, p) ? 1u : 0u;
        s_count[threadIdx.x] = n;
        __syncthreads();
        // inclusive scan of the pass's counts (Hillis-Steele)
        for (int o = 1; o < 1024; o <<= 1) {
            const uint32_t v = threadIdx.x >= (unsigned)o ? s_count[threadIdx.x - o] : 0u;
            __syncthreads();
            s_count[threadIdx.x] += v;
            __syncthreads();
        }
        const uint32_t first = base + s_count[threadIdx.x] - n;
        if (c < cells) {
            start[c] = first;
            uint32_t k = first;
            for (int p = 0; p < N; ++p)
                if (member(c, p)) list[k++] = (uint16_t)p;
        }
        base += s_count[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) start[cells] = base;
}

hipError_t launch_cell_lists(const SampleTable &st, int n, int q, float r_out, float c_base, uint32_t *start,
                             uint16_t *list, hipStream_t s)
{
    hipLaunchKernelGGL(cell_lists_kernel, dim3(1), dim3(1024), 0, s, st, n, q, r_out, c_base, start, list);
    return hipGetLastError();
}

size_t grainwise_sorted_static_lds(int block) { return (size_t)block / 64 * (kQueueCap * 4 + kOwnerMap) + 64; }

size_t grainwise_sorted_lds(int tile_w, int tile_h, int n_samples, int qbuckets)
{
    const size_t words = ((size_t)n_samples + 31) / 32;
    auto a16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
    return a16((size_t)tile_w * tile_h * words * 4) + a16((size_t)n_samples * 8) + a16((size_t)n_samples * 2) +
           a16((size_t)(qbuckets + 1) * 2);
}

hipError_t launch_grainwise_sorted(const GrainArgs &a, const SampleTable &st, int dist_kind, int block,
                                   hipStream_t s)
{
    const dim3 grid(a.n_tiles);
    const size_t lds = (size_t)a.lds_bytes;
    if (dist_kind == 0) {
        if (block == 256)
            hipLaunchKernelGGL((grainwise_sorted<0, 256>), grid, dim3(256), lds, s, a, st);
        else
            hipLaunchKernelGGL((grainwise_sorted<0, 512>), grid, dim3(512), lds, s, a, st);
    } else {
        if (block == 256)
            hipLaunchKernelGGL((grainwise_sorted<1, 256>), grid, dim3(256), lds, s, a, st);
        else
            hipLaunchKernelGGL((grainwise_sorted<1, 512>), grid, dim3(512), lds, s, a, st);
    }
    return hipGetLastError();
}

hipError_t launch_grainwise(const GrainArgs &a, int dist_kind, bool with_stats, hipStream_t s)
{
    const dim3 grid(a.n_tiles), block(kBlock);
    const size_t lds = (size_t)a.lds_bytes;
    if (dist_kind == 0) {
        if (with_stats)
            hipLaunchKernelGGL((grainwise_tiles<0, true>), grid, block, lds, s, a);
        else
            hipLaunchKernelGGL((grainwise_tiles<0, false>), grid, block, lds, s, a);
    } else {
        if (with_stats)
            hipLaunchKernelGGL((grainwise_tiles<1, true>), grid, block, lds, s, a);
        else
            hipLaunchKernelGGL((grainwise_tiles<1, false>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}

}  // namespace fg
