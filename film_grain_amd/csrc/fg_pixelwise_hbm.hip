// fg_pixelwise_hbm.hip -- pixel-wise integrator with the cell grains generated
// once per plane into HBM.
//
// Same result as the tile kernels in fg_pixelwise.hip (bit-identical; see the
// parity tests), organised so that every cell is generated exactly once:
//   threshold   per input pixel, the Poisson threshold of its cells
//               (exp(-lambda*delta*delta), or a skip / rejection marker)
//   bin_count   cells of the plane's extended grid (every cell any window can
//               reach) in 16x16-cell bins, one thread per cell: Poisson counts,
//               summed per bin
//   scan        exclusive scan of the bin counts (hipCUB) -> bin_start
//   bin_grains  each bin's grains written contiguously at bin_start[bin]: the
//               non-empty cells' counts and streams first, then the draws,
//               a lane taking the bin's next cell when its own is done
//   samples     the N sample offsets ordered by the fraction of their scaled
//               x offset, with a bucket index over that fraction
//   lists       candidate lists over that order per (radius class, cell of a
//               grain's centre fractions): the samples whose box can hold a
//               pixel column and row (splat_lists_kernel)
//   splat       per output tile, every wave streams 64-grain chunks of the
//               tile's bins on its own.  The chunk's listed grains (narrow,
//               no cell window to check) walk their lists concatenated, 64
//               entries a step whatever the grains they belong to; each lane
//               tests the pixel nearest its sample's target with the exact
//               reference test.  Every other grain walks, on its own, the one
//               cyclic run of the sample order whose x fraction can put a
//               pixel column in its candidate box: narrow ones test in the
//               lane as above, the rest queue the (grain, sample, pixel)
//               candidates whose box meets the tile and drain the queue 64 at
//               a time through the same test.  Hits go to an LDS bitset of
//               (pixel, sample) coverage.
// A grain record is (cx, cy, r, cell) with cell = (i - Ix0) | (j - Jy0) << 16.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "fg_device.h"
#include "fg_pixelwise_hbm.h"
#include "fg_region.h"

namespace fg {

// Also flags lambda values outside the accepted range (kStatusLambda), once
// per wave.
__device__ __forceinline__ void threshold_body(const float *lambda, long long n, float delta, double *T,
                                               int *status, long long p)
{
    bool bad = false;
    if (p < n) {
        const float lam = lambda[p];
        bad = lambda_invalid(lam);
        T[p] = cell_threshold(lam, delta);
    }
    if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(status, kStatusLambda);
}

__global__ __launch_bounds__(kBlock) void threshold_kernel(const float *lambda, long long n, float delta,
                                                           double *T, int *status)
{
    threshold_body(lambda, n, delta, T, status, (long long)blockIdx.x * kBlock + threadIdx.x);
}

// Cells of a bin, one wave per bin: lane l holds the four cells of column
// l & 15 and rows 4 (l >> 4) + q, q = 0..3, which share the column half of
// the cell hash (src/rng.rs:40-44 hashes i, then j).  A bin's nonempty mask
// is four 64-bit words, bit l of word q for cell (l, q).
__device__ __forceinline__ void bin_origin(const PixGeom &g, int bin, int &ci0, int &cj0)
{
    const int by = bin / g.nbx;
    ci0 = (bin - by * g.nbx) * kBinSide;
    cj0 = by * kBinSide;
}

__device__ __forceinline__ double cell_threshold_at(const PixelArgs &a, const double *T, int i, int j,
                                                    float &sx, float &sy)
{
    sx = (float)i * a.delta;
    sy = (float)j * a.delta;
    const float fx = floorf(sx), fy = floorf(sy);
    const int ix = fx < 0.0f ? 0 : (fx > (float)(a.in_w - 1) ? a.in_w - 1 : (int)fx);
    const int iy = fy < 0.0f ? 0 : (fy > (float)(a.in_h - 1) ? a.in_h - 1 : (int)fy);
    return T[(size_t)iy * a.in_w + ix];
}

// mix3 split at its two halves (src/rng.rs:40-44).
__device__ __forceinline__ uint64_t mix_col(uint64_t seed, int32_t i)
{
    return splitmix64(rotl64((seed ^ kCellStream) + (uint64_t)(int64_t)i, 17));
}
__device__ __forceinline__ uint64_t mix_row(uint64_t col, int32_t j)
{
    return splitmix64(rotl64(col + (uint64_t)(int64_t)j, 41));
}

// Poisson count of a cell known to pass the empty-cell test (or in the
// rejection branch), leaving rng after the count draws as Poisson::sample
// does (rand_distr 0.4.3: Knuth multiplies uniforms until p <= exp(-lambda)).
__device__ __forceinline__ uint32_t cell_count_full(uint64_t h, double e, Xoshiro &rng)
{
    if (e > 1.5) return 0u;
    rng.seed(h);
    if (e >= 0.0) {
        double p = standard_f64(rng.next());
        uint32_t n = 0;
        while (p > e) {
            p = p * standard_f64(rng.next());
            ++n;
        }
        return n;
    }
    const double r = poisson_rejection(rng, -e);
    if (!(r > 0.0)) return 0u;
    if (r >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)r;
}

// Writes the wave's compacted list of the bin's nonempty cells (cell (l, q)
// as q << 6 | l, in mask order) to list and returns their number.
__device__ __forceinline__ int compact_cells(const unsigned long long m[4], uint16_t *list)
{
    const int lane = threadIdx.x & 63;
    int n = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int before = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m[q] >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m[q], 0u));
        if ((m[q] >> lane) & 1ull) list[n + before] = (uint16_t)(q << 6 | lane);
        n += __popcll(m[q]);
    }
    return n;
}

// A bin's per-cell LDS records, one 32-byte record per cell in two uint4
// halves: after the empty-cell test, at the cell's slot q << 6 | l, the hash,
// two seed words and the threshold of each nonempty cell; after the counts,
// at the cell's compacted index (never above its slot, so the rewrite in
// place reads every record before it is overwritten), its stream after the
// count draws.
typedef uint4 BinRecs[2][256];

__device__ __forceinline__ void put_u64x2(uint4 &r, uint64_t x, uint64_t y)
{
    r = make_uint4((uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32));
}
__device__ __forceinline__ uint64_t lo64(const uint4 &r) { return (uint64_t)r.x | (uint64_t)r.y << 32; }
__device__ __forceinline__ uint64_t hi64(const uint4 &r) { return (uint64_t)r.z | (uint64_t)r.w << 32; }


// The grain draws of the bin's n compacted cells, whose streams are in rec
// and whose counts' exclusive prefix is in pre (total: the bin's sum): cell
// idx's grains go to grains[out0 + pre[idx] ...].  A lane takes the bin's
// next cell whenever its own has no grain left, so the lanes stay busy
// although counts differ.
template <int KIND>
__device__ __forceinline__ void bin_draws(const PixelArgs &a, const PixGeom &g, int ci0, int cj0,
                                          const uint16_t *list, int n, const BinRecs &rec, const uint32_t *pre,
                                          uint32_t total, uint32_t out0, float4 *grains)
{
    Xoshiro rng{};
    uint32_t k = 0, kend = 0, dst = 0;
    float sx = 0.0f, sy = 0.0f, pkf = 0.0f;
    int next = 0;  // uniform: the bin's next cell no lane has taken
    for (;;) {
        while (next < n) {
            const bool need = k >= kend;
            const unsigned long long mk = __ballot(need);
            if (mk == 0) break;
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u));
            const int idx = next + rank;
            if (need && idx < n) {
                const int cell = list[idx];
                const int l = cell & 63, q = cell >> 6;
                const int ci = ci0 + (l & 15), cj = cj0 + 4 * (l >> 4) + q;
                const uint32_t p0 = pre[idx];
                const uint32_t p1 = idx + 1 < n ? pre[idx + 1] : total;
                const uint4 r0 = rec[0][idx], r1 = rec[1][idx];
                rng.s0 = lo64(r0);
                rng.s1 = hi64(r0);
                rng.s2 = lo64(r1);
                rng.s3 = hi64(r1);
                k = 0;
                kend = p1 - p0;
                dst = out0 + p0;
                sx = (float)(g.Ix0 + ci) * a.delta;
                sy = (float)(g.Jy0 + cj) * a.delta;
                pkf = __uint_as_float((uint32_t)ci | ((uint32_t)cj << 16));
            }
            next = min(next + __popcll(mk), n);
        }
        if (!__any(k < kend)) break;
        if (k < kend) {
            const float cx = sx + uniform_f32(rng, a.uscale);
            const float cy = sy + uniform_f32(rng, a.uscale);
            float rad = a.r_const;
            if (KIND == 1) {
                rad = lognormal_f32(rng, a.log_mu, a.log_sigma, d_zig_x, d_zig_f);
                if (rad > a.rm) rad = a.rm;
            }
            grains[dst + k] = make_float4(cx, cy, rad, pkf);
            ++k;
        }
    }
}

// Count of a rejection-branch cell (lambda * delta^2 >= 12) from its seeded
// stream, saturated as cell_count_full does.  Kept out of line: few cells take
// it, and inlined its f64 transcendentals would hold registers across the
// whole counting loop of bin_count_planes_kernel.
__device__ __noinline__ uint32_t rejection_count(Xoshiro rng, double lambda)
{
    const double r = poisson_rejection(rng, lambda);
    return !(r > 0.0) ? 0u : (r >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)r);
}

// Pass 1 over the extended cell grid, per bin (one wave): each of its 256
// cells' empty-cell test, the nonempty masks and the bin's grain count, for
// the NP <= kFramePlanes planes of one frame (NP = 1 for a single plane).
// A cell is empty when its Poisson count is 0 by the Knuth branch's first
// draw (u1 <= exp(-lambda)), which needs only the two seed words that draw
// depends on, s0 and s3; the rejection branch (e < 0) and the all-zero seed
// case (reseeded by seed_from_u64) are counted in full.  The reference
// renders every plane with the same seed, and a cell's stream depends only
// on the seed and the cell (src/rng.rs:26-29, src/lib.rs:146-153), so the
// planes share each cell's hash, seed words and Knuth uniforms; only their
// thresholds differ.  The hashes and the test's two seed words are computed
// once per cell and kept in an LDS record for the nonempty ones, and each
// plane's count of a cell runs on the one uniform sequence against its own
// threshold.  A bin of 2^32 or more grains saturates its count and sets
// kStatusOverflow (the plane's 64-bit sum of the bin counts, sum_bins_u64,
// then reaches 2^32 - 1 as well, and the host refuses the plane before any
// kernel uses the u32 scan).
// A nonempty cell's record in the count pass: its hash and seed words s0
// (16 B) and s3 (8 B); 24 B a cell keep a workgroup's LDS under 32 KB.
struct CountRecs {
    uint4 hs0[256];
    uint2 s3[256];
};

// One plane runs five waves per SIMD (100 VGPRs, 26 KB of LDS a
// workgroup); more planes need the registers of four.
template <int NP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(NP == 1 ? 5 : 4, 8))) void
bin_count_planes_kernel(PixelArgs a, PixGeom g, FramePlanes fp)
{
    constexpr int NW = kBlock / 64;
    __shared__ uint16_t s_list[NW][256];
    __shared__ CountRecs s_rec[NW];
    __shared__ unsigned long long s_mask[NW][NP + 1][4];  // the planes' nonempty masks, then their union
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int bin = blockIdx.x * NW + wid;
    if (bin >= g.nbx * g.nby) return;
    int ci0, cj0;
    bin_origin(g, bin, ci0, cj0);
    const int ci = ci0 + (lane & 15), cjl = cj0 + 4 * (lane >> 4);
    const uint64_t col = mix_col(a.seed, g.Ix0 + ci);
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const int cj = cjl + q;
        bool ne[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) ne[p] = false;
        if (ci < g.W && cj < g.H) {
            double e[NP];
            bool knuth = false;
            float sx, sy;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                e[p] = cell_threshold_at(a, fp.T[p], g.Ix0 + ci, g.Jy0 + cj, sx, sy);
                knuth |= e[p] >= 0.0 && e[p] <= 1.5;
            }
            const uint64_t h = mix_row(col, g.Jy0 + cj);
            uint64_t s0 = 0, s3 = 0;
            double u1 = 0.0;
            if (knuth) {
                s0 = splitmix64(h);
                s3 = splitmix64(h + 3 * kGolden);
                u1 = standard_f64(rotl64(s0 + s3, 23) + s0);
            }
            bool any = false;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                // the rejection branch and the all-zero seed case are
                // counted in full
                ne[p] = e[p] <= 1.5 && (e[p] < 0.0 || (s0 | s3) == 0 || u1 > e[p]);
                any |= ne[p];
            }
            if (any) {
                const int slot = q << 6 | lane;
                put_u64x2(s_rec[wid].hs0[slot], h, s0);
                s_rec[wid].s3[slot] = make_uint2((uint32_t)s3, (uint32_t)(s3 >> 32));
            }
        }
        unsigned long long any = 0ull;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const unsigned long long mp = __ballot(ne[p]);
            if (lane == 0) s_mask[wid][p][q] = mp;
            any |= mp;
        }
        if (lane == 0) s_mask[wid][NP][q] = any;
    }
    unsigned long long mu[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) mu[q] = s_mask[wid][NP][q];
    const int n = compact_cells(mu, s_list[wid]);
    unsigned long long total[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) total[p] = 0;
    for (int base = 0; base < n; base += 64) {
        uint32_t c[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) c[p] = 0;
        if (base + lane < n) {
            const int slot = s_list[wid][base + lane];
            const int l = slot & 63, q = slot >> 6;
            const uint4 r0 = s_rec[wid].hs0[slot];
            const uint2 r1 = s_rec[wid].s3[slot];
            const uint64_t h = lo64(r0);
            const int i = g.Ix0 + ci0 + (l & 15), j = g.Jy0 + cj0 + 4 * (l >> 4) + q;
            double e[NP];
            bool on[NP], knuth = false;
            float sx, sy;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                on[p] = (s_mask[wid][p][q] >> l) & 1ull;  // this plane's bit of the cell
                e[p] = cell_threshold_at(a, fp.T[p], i, j, sx, sy);
                knuth |= on[p] && e[p] >= 0.0;
            }
            Xoshiro seeded;
            seeded.s0 = hi64(r0);
            seeded.s1 = splitmix64(h + kGolden);
            seeded.s2 = splitmix64(h + 2 * kGolden);
            seeded.s3 = (uint64_t)r1.x | (uint64_t)r1.y << 32;
            // a cell counted in full for the rejection branch alone has no
            // recorded seed words
            if (!knuth || (seeded.s0 | seeded.s1 | seeded.s2 | seeded.s3) == 0) seeded.seed(h);
            if (knuth) {
                // Knuth: count = multiplications until the product <= e
                Xoshiro rng = seeded;
                double pr = standard_f64(rng.next());
                uint32_t k = 0;
                bool pend[NP], left = false;
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    pend[p] = on[p] && e[p] >= 0.0 && pr > e[p];
                    left |= pend[p];
                }
                while (left) {
                    pr = pr * standard_f64(rng.next());
                    ++k;
                    left = false;
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        if (pend[p] && !(pr > e[p])) {
                            c[p] = k;
                            pend[p] = false;
                        }
                        left |= pend[p];
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                if (on[p] && e[p] < 0.0) c[p] = rejection_count(seeded, -e[p]);
            }
        }
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            unsigned long long x = c[p];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            total[p] += x;
        }
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        if (lane == 0) {
            fp.cnt[p][bin] = total[p] > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)total[p];
            if (total[p] > 0xFFFFFFFFull) atomicOr(a.status, kStatusOverflow);
        }
        if (lane < 4) fp.mask[p][(size_t)bin * 4 + lane] = s_mask[wid][p][lane];
    }
}

// Pass 2: per bin (one wave), the nonempty cells from the mask, their counts
// (seeded in full) and their grains at bin_start[bin].
template <int KIND>
__device__ __forceinline__ void bin_grains_body(const PixelArgs &a, const PixGeom &g, const double *T,
                                                const uint32_t *bin_start, const unsigned long long *bin_mask,
                                                float4 *grains, int bin, uint16_t *list, uint32_t *pre, BinRecs &rec)
{
    const int lane = threadIdx.x & 63;
    unsigned long long m[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) m[q] = bin_mask[(size_t)bin * 4 + q];
    const int n = compact_cells(m, list);
    int ci0, cj0;
    bin_origin(g, bin, ci0, cj0);
    // counts from full seeds
    uint32_t carry = 0;
    for (int base = 0; base < n; base += 64) {
        uint32_t c = 0;
        const int idx = base + lane;
        if (idx < n) {
            const int cell = list[idx];
            const int l = cell & 63, q = cell >> 6;
            const int i = g.Ix0 + ci0 + (l & 15), j = g.Jy0 + cj0 + 4 * (l >> 4) + q;
            float sx, sy;
            const double e = cell_threshold_at(a, T, i, j, sx, sy);
            Xoshiro rng{};
            c = cell_count_full(mix3(a.seed, kCellStream, i, j), e, rng);
            put_u64x2(rec[0][idx], rng.s0, rng.s1);
            put_u64x2(rec[1][idx], rng.s2, rng.s3);
        }
        const uint32_t incl = wave_inclusive_sum(c);
        if (idx < n) pre[idx] = carry + incl - c;
        carry += __shfl(incl, 63, 64);
    }
    bin_draws<KIND>(a, g, ci0, cj0, list, n, rec, pre, carry, bin_start[bin], grains);
}

template <int KIND>
__global__ __launch_bounds__(kBlock, 4) void bin_grains_kernel(PixelArgs a, PixGeom g, const double *T,
                                                               const uint32_t *bin_start,
                                                               const unsigned long long *bin_mask, float4 *grains)
{
    constexpr int NW = kBlock / 64;
    __shared__ uint16_t s_list[NW][256];
    __shared__ uint32_t s_pre[NW][256];
    __shared__ BinRecs s_rec[NW];
    const int wid = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    if (a.grain_total && *a.grain_total > a.grain_cap) return;
    const int b0 = (blockIdx.x * NW + wid) * kDrawChunk, b1 = min(g.nbx * g.nby, b0 + kDrawChunk);
#pragma unroll 1
    for (int bin = b0; bin < b1; ++bin)
        bin_grains_body<KIND>(a, g, T, bin_start, bin_mask, grains, bin, s_list[wid], s_pre[wid], s_rec[wid]);
}

// The batch's generation kernels run over all its planes in one launch each:
// a workgroup or wave finds its plane (or group of planes) from its index in
// the launch by a binary search over first-index offsets.  The offsets and
// plane records are read through the constant address space: the index is
// uniform, so these are scalar loads into SGPRs.
template <class T>
__device__ __forceinline__ T cload(const T *p)
{
    return *(const T *)((const __attribute__((address_space(4))) T *)p);
}

__device__ __forceinline__ const BatchPlane &cplane(const BatchPlane *planes, int p)
{
    return *(const BatchPlane *)((const __attribute__((address_space(4))) BatchPlane *)planes + p);
}

// The run i with off[i] <= x < off[i + 1], for 0 <= x < off[n] (empty runs
// are passed over).
__device__ __forceinline__ int find_run(const long long *off, int n, long long x)
{
    int lo = 0, hi = n;  // off[lo] <= x < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cload(off + mid) <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void batch_threshold_kernel(BatchGen bg)
{
    const long long blk = blockIdx.x;
    const int p = find_run(bg.thr_off, bg.n, blk);
    const BatchPlane &bp = cplane(bg.planes, p);
    const long long i = (blk - cload(bg.thr_off + p)) * kBlock + threadIdx.x;
    threshold_body(bp.a.lambda, (long long)bp.a.in_w * bp.a.in_h, bp.a.delta, cload(bg.T + p), bp.a.status, i);
}

template <int KIND>
__global__ __launch_bounds__(kBlock, 4) void batch_bin_grains_kernel(BatchGen bg)
{
    constexpr int NW = kBlock / 64;
    __shared__ uint16_t s_list[NW][256];
    __shared__ uint32_t s_pre[NW][256];
    __shared__ BinRecs s_rec[NW];
    const int wid = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const long long gc = (long long)blockIdx.x * NW + wid;
    const long long *off = bg.kind_off[KIND];
    const int nk = bg.n_kind[KIND];
    if (gc >= cload(off + nk)) return;
    const int i = find_run(off, nk, gc);
    const int p = cload(bg.kind_plane[KIND] + i);
    const BatchPlane &bp = cplane(bg.planes, p);
    if (bp.a.grain_total && *bp.a.grain_total > bp.a.grain_cap) return;
    const int b0 = (int)(gc - cload(off + i)) * kDrawChunk;
    const int b1 = min(bp.geo.nbx * bp.geo.nby, b0 + kDrawChunk);
    const double *T = cload(bg.T + p);
    const unsigned long long *mask = bg.mask + 4 * cload(bg.bin_off + p);
#pragma unroll 1
    for (int bin = b0; bin < b1; ++bin)
        bin_grains_body<KIND>(bp.a, bp.geo, T, bp.bin_start, mask, bg.grains, bin, s_list[wid], s_pre[wid],
                              s_rec[wid]);
}

// ---------------------------------------------------------------------------
// sample order
// ---------------------------------------------------------------------------
// Fraction of a scaled offset in [0, 1); 1.0 (rounded up from just below) and
// NaN map to 0.  The sample order and the grain queries only need the two to
// agree up to the query's widening, which covers this rounding.
__device__ __forceinline__ float sample_frac(float X)
{
    const float f = X - floorf(X);
    return f < 1.0f ? f : 0.0f;
}

// Rank sort of the samples by (fraction, index): thread i writes sample i to
// its rank.  O(N^2) compares, N <= 65535, run once per render.
__device__ __forceinline__ void sample_sort_body(const float2 *offs, int n, float zoom, const SampleTable &t,
                                                 int blk, float *s_key)
{
    const int i = blk * kBlock + threadIdx.x;
    float2 oi = make_float2(0.0f, 0.0f);
    float f = 0.0f;
    if (i < n) {
        oi = offs[i];
        f = sample_frac(oi.x * zoom);
    }
    int rank = 0;
    for (int base = 0; base < n; base += kBlock) {
        const int j = base + (int)threadIdx.x;
        __syncthreads();
        s_key[threadIdx.x] = j < n ? sample_frac(offs[j].x * zoom) : 0.0f;
        __syncthreads();
        const int m = min(kBlock, n - base);
        for (int jj = 0; jj < m; ++jj) {
            const float fj = s_key[jj];
            rank += (fj < f || (fj == f && base + jj < i)) ? 1 : 0;
        }
    }
    if (i < n) {
        const_cast<float2 *>(t.xy)[rank] = make_float2(oi.x * zoom, oi.y * zoom);
        const_cast<float2 *>(t.o)[rank] = oi;
        const_cast<uint16_t *>(t.k)[rank] = (uint16_t)i;
        const_cast<float *>(t.key)[rank] = f;
    }
}

__global__ __launch_bounds__(kBlock) void sample_sort_kernel(const float2 *offs, int n, float zoom,
                                                             SampleTable t)
{
    __shared__ float s_key[kBlock];
    sample_sort_body(offs, n, zoom, t, blockIdx.x, s_key);
}

// kRowBuckets intervals over a table's scaled y offsets [ext.z, ext.w]:
// bucket(y) = floor((y - ext.z) * inv_w) clamped to [0, kRowBuckets - 1].
// It is monotone in y, so the samples with y in [lo, hi] lie in buckets
// bucket(lo)..bucket(hi).  sample_index (which orders the table) and the
// splat compute it with the same code from the same extent; the splat's
// bounds are widened by 2^-4 besides, far more than a last-bit difference.
struct RowBuckets {
    float y0, inv_w;
};
__device__ __forceinline__ RowBuckets row_buckets(const float4 &ext)
{
    RowBuckets r;
    r.y0 = ext.z;
    r.inv_w = (float)kRowBuckets / ((ext.w - ext.z) + 0.015625f);
    return r;
}
__device__ __forceinline__ int row_bucket(const RowBuckets &r, float y)
{
    return min(max((int)floorf((y - r.y0) * r.inv_w), 0), kRowBuckets - 1);
}

// Bucket index (first sorted position with fraction >= q / Q, by binary
// search), the extents of the scaled offsets, and the (X.y, position) order
// with its row-bucket starts; one workgroup.
struct IndexShared {
    float4 r[kBlock];
    float y[2048];
    uint16_t ord[2048];
};

__device__ __forceinline__ void sample_index_body(int n, const SampleTable &t, IndexShared &sh)
{
    float4 *s_r = sh.r;
    const int Q = t.qbuckets;
    for (int q = threadIdx.x; q <= Q; q += kBlock) {
        const float v = (float)q / (float)Q;  // exact: Q is a power of two
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (t.key[mid] < v)
                lo = mid + 1;
            else
                hi = mid;
        }
        const_cast<uint16_t *>(t.qidx)[q] = (uint16_t)lo;
    }
    float4 r = make_float4(INFINITY, -INFINITY, INFINITY, -INFINITY);
    for (int k = threadIdx.x; k < n; k += kBlock) {
        const float2 X = t.xy[k];
        r.x = fminf(r.x, X.x);
        r.y = fmaxf(r.y, X.x);
        r.z = fminf(r.z, X.y);
        r.w = fmaxf(r.w, X.y);
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
    if (threadIdx.x == 0) *const_cast<float4 *>(t.ext) = s_r[0];
    // the (X.y, position) order by rank (n <= 2048 for these; larger tables
    // have no row order and leave it unused)
    if (n > 2048) return;
    const RowBuckets rb = row_buckets(s_r[0]);
    for (int k = threadIdx.x; k < n; k += kBlock) sh.y[k] = t.xy[k].y;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += kBlock) {
        const float y = sh.y[k];
        int r = 0;
        for (int j = 0; j < n; ++j) {
            const float yj = sh.y[j];
            r += (yj < y || (yj == y && j < k)) ? 1 : 0;
        }
        sh.ord[r] = (uint16_t)k;
    }
    __syncthreads();
    uint16_t *yord = const_cast<uint16_t *>(t.yord), *ybk = const_cast<uint16_t *>(t.ybk);
    for (int i = threadIdx.x; i < n; i += kBlock) yord[i] = sh.ord[i];
    // the row buckets of the order are non-decreasing; index i starts
    // buckets (bucket(i - 1), bucket(i)], with bucket(-1) = 0 and
    // bucket(n) = kRowBuckets
    if (threadIdx.x == 0) ybk[0] = 0;
    for (int i = threadIdx.x; i <= n; i += kBlock) {
        const int prev = i == 0 ? 0 : row_bucket(rb, sh.y[sh.ord[i - 1]]);
        const int cur = i == n ? kRowBuckets : row_bucket(rb, sh.y[sh.ord[i]]);
        for (int b = prev + 1; b <= cur; ++b) ybk[b] = (uint16_t)i;
    }
}

__global__ __launch_bounds__(kBlock) void sample_index_kernel(int n, SampleTable t)
{
    __shared__ IndexShared sh;
    sample_index_body(n, t, sh);
}

// The table of the first n samples from the table of all of them: the full
// order is sorted by (fraction, index), so keeping the entries with index < n
// in that order sorts the prefix by the same key.  One workgroup of 1024
// threads, full.n <= 2048.
__global__ __launch_bounds__(1024) void sample_prefix_kernel(SampleTable full, int n, SampleTable t)
{
    __shared__ uint32_t s_scan[1024];
    const int i0 = 2 * (int)threadIdx.x, N = full.n;
    const uint32_t f0 = (i0 < N && full.k[i0] < n) ? 1u : 0u;
    const uint32_t f1 = (i0 + 1 < N && full.k[i0 + 1] < n) ? 1u : 0u;
    s_scan[threadIdx.x] = f0 + f1;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t v = (int)threadIdx.x >= o ? s_scan[threadIdx.x - o] : 0u;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t pos = s_scan[threadIdx.x] - (f0 + f1);
    for (int j = 0; j < 2; ++j) {
        const int i = i0 + j;
        if (j == 0 ? f0 : f1) {
            const_cast<float2 *>(t.xy)[pos] = full.xy[i];
            const_cast<float2 *>(t.o)[pos] = full.o[i];
            const_cast<uint16_t *>(t.k)[pos] = full.k[i];
            const_cast<float *>(t.key)[pos] = full.key[i];
            ++pos;
        }
    }
}

__host__ __device__ static inline size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

size_t sample_table_bytes(int n, int qbuckets)
{
    return 16 + align16((size_t)n * 8) * 2 + align16((size_t)n * 4) + align16((size_t)n * 2) * 2 +
           align16((size_t)(qbuckets + 1) * 2) + align16((size_t)(kRowBuckets + 1) * 2);
}

hipError_t launch_sample_table(const float2 *offs, int n, float zoom, int qbuckets, void *buf, SampleTable *t,
                               hipStream_t s)
{
    sample_table_layout(buf, n, qbuckets, t);
    hipLaunchKernelGGL(sample_sort_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, offs,
                       n, zoom, *t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sample_index_kernel, dim3(1), dim3(kBlock), 0, s, n, *t);
    return hipGetLastError();
}

hipError_t launch_sample_prefix(const SampleTable &full, int n, int qbuckets, void *buf, SampleTable *t,
                                hipStream_t s)
{
    if (full.n > 2048 || n > full.n) return hipErrorInvalidValue;
    sample_table_layout(buf, n, qbuckets, t);
    hipLaunchKernelGGL(sample_prefix_kernel, dim3(1), dim3(1024), 0, s, full, n, *t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sample_index_kernel, dim3(1), dim3(kBlock), 0, s, n, *t);
    return hipGetLastError();
}

void sample_table_layout(void *buf, int n, int qbuckets, SampleTable *t)
{
    unsigned char *p = static_cast<unsigned char *>(buf);
    t->ext = reinterpret_cast<float4 *>(p);
    p += 16;
    t->xy = reinterpret_cast<float2 *>(p);
    p += align16((size_t)n * 8);
    t->o = reinterpret_cast<float2 *>(p);
    p += align16((size_t)n * 8);
    t->key = reinterpret_cast<float *>(p);
    p += align16((size_t)n * 4);
    t->k = reinterpret_cast<uint16_t *>(p);
    p += align16((size_t)n * 2);
    t->qidx = reinterpret_cast<uint16_t *>(p);
    p += align16((size_t)(qbuckets + 1) * 2);
    t->yord = reinterpret_cast<uint16_t *>(p);
    p += align16((size_t)n * 2);
    t->ybk = reinterpret_cast<uint16_t *>(p);
    t->n = n;
    t->qbuckets = qbuckets;
}

// ---------------------------------------------------------------------------
// splat from HBM grains
// ---------------------------------------------------------------------------
// Candidate lists for the direct test.  A grain's candidate box has a column
// only if frac(X) lies within rz + eps of c = frac(-bx), bx = cx * zoom - 0.5
// (c is exact in f32, and so is its cell floor(c * q) for q a power of two):
// the f32 box and cxk = bx + X take a few roundings of at most C * 2^-24
// each, which eps covers.  Likewise rows with frac(Y) and frac(-by).  The
// list of class k (rz < R = (k + 1) dz) and cell (qa, qb) holds, in sorted
// order, every sample with frac(X) in [qa/q - R - eps, (qa+1)/q + R + eps]
// and frac(Y) in [qb/q - R - eps, (qb+1)/q + R + eps] (mod 1), each widened
// by 2^-12 for the rounding of this test and of the class.  With dz =
// 1 / (2 q) a window is (k + 2 + (2 eps + 2^-11) q) / q wide, so a sample
// lies in at most splat_list_span(k, q, e) cells per axis when e >=
// (2 eps + 2^-11) q, and class k needs at most n * span^2 list entries.
// An entry is its sample's sorted position p times 16, the byte offset of
// the sample's record in the splat's LDS (n <= 2048 keeps it in 16 bits).
// A list holds its entries in order of the scaled y offset X.y (ties by p),
// with per-cell counts by row bucket (row_bucket), so that the splat walks
// only the entries whose target row can lie in its tile.
// Blocks of 64 cells of one class, four threads per cell; both passes test
// every sample, and each block reserves its cells' room in the class's
// region with one atomic.
__host__ __device__ inline uint32_t splat_list_span(int k, int q, int e)
{
    return (uint32_t)min(q, k + 4 + e);
}

__host__ __device__ inline size_t splat_list_base(int k, int q, int e, int n)
{
    size_t b = 0;
    for (int j = 0; j < k; ++j) b += (size_t)n * splat_list_span(j, q, e) * splat_list_span(j, q, e);
    return b;
}

struct ListsShared {
    float2 f[2048];       // frac(X) of sorted position p
    uint16_t ord[2048];   // the table's (X.y, position) order
    int jb[kRowBuckets + 1];  // jb[b]: first index of that order in row bucket >= b
    uint32_t scan[256];
    uint32_t base;
    bool fits;
};

__device__ __forceinline__ void splat_lists_body(const SampleTable &st, int q, int e, uint32_t *fill, uint2 *cell,
                                                 uint16_t *yidx, uint16_t *list, int *status, int blk,
                                                 ListsShared &sh)
{
    // 64 cells per block, each split over 4 adjacent threads that take the
    // four quarters of the (X.y, p) order: the scan over the block in thread
    // order keeps every list in that order
    float2 *s_f = sh.f;
    uint32_t *s_scan = sh.scan;
    uint32_t &s_base = sh.base;
    bool &s_fits = sh.fits;
    const int cells = q * q, parts = (cells + 63) / 64, N = st.n;
    const int k = blk / parts, c = (blk - k * parts) * 64 + ((int)threadIdx.x >> 2);
    const int quarter = threadIdx.x & 3, qn = (N + 3) / 4;
    const int p_lo = quarter * qn, p_hi = min(N, p_lo + qn);
    for (int p = threadIdx.x; p < N; p += 256) {
        const float2 X = st.xy[p];
        s_f[p] = make_float2(X.x - floorf(X.x), X.y - floorf(X.y));
        sh.ord[p] = st.yord[p];
    }
    if (threadIdx.x <= kRowBuckets) sh.jb[threadIdx.x] = st.ybk[threadIdx.x];
    __syncthreads();
    const float m = 2.44140625e-4f;  // 2^-12
    const float R = (float)(k + 1) / (float)(2 * q);
    const float w = ((R + R) + (st.eps + st.eps)) + (1.0f / (float)q + (m + m));
    const float lx = ((float)(c % q) / (float)q - R) - (st.eps + m);
    const float ly = ((float)(c / q) / (float)q - R) - (st.eps + m);
    auto member = [&](int p) {
        const float2 f = s_f[p];
        float dx = f.x - lx, dy = f.y - ly;
        dx -= floorf(dx);
        dy -= floorf(dy);
        return dx <= w && dy <= w;
    };
    uint32_t n = 0;
    if (c < cells)
        for (int i = p_lo; i < p_hi; ++i) n += member(sh.ord[i]) ? 1u : 0u;
    // the block's cells take one contiguous range of the class's room
    s_scan[threadIdx.x] = n;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t v = (int)threadIdx.x >= o ? s_scan[threadIdx.x - o] : 0u;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    if (threadIdx.x == 255) {
        const uint32_t span = splat_list_span(k, q, e);
        const uint32_t off = atomicAdd(&fill[k], s_scan[255]);
        s_base = (uint32_t)splat_list_base(k, q, e, N) + off;
        // the room argument above; a block beyond it writes nothing and
        // fails the render instead
        s_fits = (size_t)off + s_scan[255] <= (size_t)N * span * span;
        if (!s_fits) atomicOr(status, kStatusCapacity);
    }
    __syncthreads();
    if (c < cells && s_fits) {
        uint32_t pos = s_base + s_scan[threadIdx.x] - n;
        // the cell's entries before this quarter
        const int t0 = (int)threadIdx.x - quarter;
        uint32_t rel = s_scan[threadIdx.x] - n - (t0 > 0 ? s_scan[t0 - 1] : 0u);
        uint16_t *yx = yidx + ((size_t)k * cells + c) * kRowBuckets;
        if (quarter == 0) {
            const uint32_t len = s_scan[threadIdx.x + 3] - (s_scan[threadIdx.x] - n);
            cell[(size_t)k * cells + c] = make_uint2(pos, len);
            // buckets that start past the last sample hold no entry
            for (int b = 1; b <= kRowBuckets; ++b)
                if (sh.jb[b] == N) yx[b - 1] = (uint16_t)len;
        }
        // each other bucket count is written by the quarter holding the
        // bucket's first index: the cell's entries before that index
        int b = 1;
        while (b <= kRowBuckets && sh.jb[b] < p_lo) ++b;
        for (int i = p_lo; i < p_hi; ++i) {
            while (b <= kRowBuckets && sh.jb[b] <= i) {
                yx[b - 1] = (uint16_t)rel;
                ++b;
            }
            const int p = sh.ord[i];
            if (member(p)) {
                list[pos++] = (uint16_t)(p << 4);
                ++rel;
            }
        }
    }
}

__global__ __launch_bounds__(256) void splat_lists_kernel(SampleTable st, int q, int e, uint32_t *fill, uint2 *cell,
                                                          uint16_t *yidx, uint16_t *list, int *status)
{
    __shared__ ListsShared sh;
    splat_lists_body(st, q, e, fill, cell, yidx, list, status, blockIdx.x, sh);
}

// Batch sample tables: workgroup (x, p) sorts samples x * kBlock.. of plane p.
__global__ __launch_bounds__(kBlock) void batch_sample_sort_kernel(BatchGen bg)
{
    __shared__ float s_key[kBlock];
    const BatchPlane &bp = cplane(bg.planes, blockIdx.y);
    if ((int)blockIdx.x * kBlock >= bp.a.n_samples) return;
    sample_sort_body(bp.a.offs, bp.a.n_samples, bp.a.zoom, bp.st, blockIdx.x, s_key);
}

// One workgroup per plane: the table's index, and the candidate lists'
// counters cleared for batch_splat_lists_kernel.
__global__ __launch_bounds__(kBlock) void batch_sample_index_kernel(BatchGen bg)
{
    __shared__ IndexShared sh;
    const int p = blockIdx.x;
    const BatchPlane &bp = cplane(bg.planes, p);
    sample_index_body(bp.a.n_samples, bp.st, sh);
    const int classes = cload(bg.list_plan + p).y;
    uint32_t *fill = cload(bg.list_fill + p);
    for (int k = threadIdx.x; k < classes; k += kBlock) fill[k] = 0u;
}

__global__ __launch_bounds__(256) void batch_splat_lists_kernel(BatchGen bg)
{
    __shared__ ListsShared sh;
    const long long blk = blockIdx.x;
    const int p = find_run(bg.list_off, bg.n, blk);
    const BatchPlane &bp = cplane(bg.planes, p);
    const int4 lp = cload(bg.list_plan + p);
    splat_lists_body(bp.st, lp.x, lp.z, cload(bg.list_fill + p), const_cast<uint2 *>(bp.st.cl_cell),
                     const_cast<uint16_t *>(bp.st.cl_yidx), const_cast<uint16_t *>(bp.st.cl_list), bp.a.status,
                     (int)(blk - cload(bg.list_off + p)), sh);
}

__global__ void store_table_kernel(SampleTable t, SampleTable *dst)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = t;
}

hipError_t launch_store_table(const SampleTable &t, SampleTable *dst, hipStream_t s)
{
    hipLaunchKernelGGL(store_table_kernel, dim3(1), dim3(64), 0, s, t, dst);
    return hipGetLastError();
}

size_t splat_lists_entries(int q, int classes, int e, int n)
{
    return splat_list_base(classes, q, e, n);
}

hipError_t launch_splat_lists(const SampleTable &st, int q, int classes, int e, uint32_t *fill, uint2 *cell,
                              uint16_t *yidx, uint16_t *list, int *status, hipStream_t s)
{
    if (q < 1 || q * q > 65536 || classes < 1 || st.n > 2048) return hipErrorInvalidValue;
    // a kernel, not hipMemsetAsync: batches capture this into a graph
    hipError_t err = launch_fill_u32(fill, classes, 0u, s);
    if (err != hipSuccess) return err;
    const int parts = (q * q + 63) / 64;
    hipLaunchKernelGGL(splat_lists_kernel, dim3(classes * parts), dim3(256), 0, s, st, q, e, fill, cell, yidx,
                       list, status);
    return hipGetLastError();
}

// Steps of the splat's concatenated list walk whose list loads are in flight
// together (4: 2% slower at C2, 12: 2% slower, 16: spills; with the lists
// cut to the tile's rows, 4 and 6 are 1.8% and 0.1% slower on a 4K RGB
// frame and 1.1% and 0.3% on a batch of 64 uneven images).
constexpr int kWalkStage = 8;

size_t splat_hbm_lds(int tile_w, int tile_h, int n_samples, int qbuckets)
{
    const size_t words = ((size_t)n_samples + 31) / 32;
    return align16((size_t)tile_w * tile_h * words * 4) + align16((size_t)n_samples * 16) +
           align16((size_t)(qbuckets + 1) * 2);
}


// Work of each splat tile, the grains of the bins covering its region, for
// a heaviest-first launch order.
__global__ __launch_bounds__(kBlock) void pixel_tile_weights_kernel(PixelArgs a, PixGeom geo,
                                                                    const uint32_t *bin_start, uint32_t *w, int *ids)
{
    const int tile = blockIdx.x * kBlock + threadIdx.x;
    if (tile >= a.n_tiles) return;
    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    const int X0 = txi * a.tile_w, Y0 = tyi * a.tile_h;
    const TileRegion r = tile_region(a, geo, X0, Y0, min(X0 + a.tile_w, a.out_w), min(Y0 + a.tile_h, a.out_h));
    uint32_t sum = 0;
    if (r.bx0 <= r.bx1)
        for (int by = r.by0; by <= r.by1; ++by)
            sum += bin_start[by * geo.nbx + r.bx1 + 1] - bin_start[by * geo.nbx + r.bx0];
    if (a.budget) {  // adaptive: the work also scales with the tile's samples
        const int n = reinterpret_cast<const SampleTable *>(a.level_tables)[a.tile_level[tile]].n;
        sum = (uint32_t)min((unsigned long long)sum * (unsigned long long)n / 16ull, 0xFFFFFFFFull);
    }
    w[tile] = sum;
    ids[tile] = tile;
}

// Adaptive sampling: per tile, the first level whose sample count covers the
// largest budget among its pixels (levels ascending, the last = n_samples).
__global__ __launch_bounds__(kBlock) void tile_levels_kernel(PixelArgs a, const int *level_n, int n_levels,
                                                             uint8_t *tile_level)
{
    const int tile = blockIdx.x * kBlock + threadIdx.x;
    if (tile >= a.n_tiles) return;
    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    const int X0 = txi * a.tile_w, Y0 = tyi * a.tile_h;
    const int X1 = min(X0 + a.tile_w, a.out_w), Y1 = min(Y0 + a.tile_h, a.out_h);
    int m = 1;
    for (int y = Y0; y < Y1; ++y)
        for (int x = X0; x < X1; ++x) m = max(m, (int)a.budget[(size_t)y * a.out_w + x]);
    int l = 0;
    while (l + 1 < n_levels && level_n[l] < m) ++l;
    tile_level[tile] = (uint8_t)l;
}

hipError_t launch_tile_levels(const PixelArgs &a, const int *level_n, int n_levels, uint8_t *tile_level,
                              hipStream_t s)
{
    hipLaunchKernelGGL(tile_levels_kernel, dim3((unsigned)((a.n_tiles + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       a, level_n, n_levels, tile_level);
    return hipGetLastError();
}

hipError_t launch_pixel_tile_weights(const PixelArgs &a, const PixGeom &g, const uint32_t *bin_start, uint32_t *w,
                                     int *ids, hipStream_t s)
{
    hipLaunchKernelGGL(pixel_tile_weights_kernel, dim3((unsigned)((a.n_tiles + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, a, g, bin_start, w, ids);
    return hipGetLastError();
}

hipError_t sort_tiles(void *temp, size_t *temp_bytes, const uint32_t *w_in, uint32_t *w_out, const int *ids_in,
                      int *ids_out, int n, hipStream_t s)
{
    return hipcub::DeviceRadixSort::SortPairsDescending(temp, *temp_bytes, w_in, w_out, ids_in, ids_out, n, 0, 32,
                                                        s);
}

__device__ __forceinline__ float lane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ int lane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// LDS locations as 32-bit LDS addresses: a bitset word's address held in a
// sample record is used as it is, with no workgroup base added per hit
typedef __attribute__((address_space(3))) uint32_t lds_u32;
__device__ __forceinline__ uint32_t lds_addr(uint32_t *p)
{
    return (uint32_t)(uintptr_t)(lds_u32 *)p;
}
__device__ __forceinline__ void lds_or(uint32_t addr, uint32_t v)
{
    __hip_atomic_fetch_or((lds_u32 *)(uintptr_t)addr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// a * b + c for a, b < 2^24 with one full-rate v_mad_u32_u24 (b uniform).
// Written out because the compiler turns two such terms into two products
// and a three-way add.
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t d;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
    return d;
}

// One output tile of one plane: the body of the single-plane kernel and of the
// persistent batch kernel (which therefore give identical planes).  lds is the
// workgroup's dynamic LDS (a.lds_bytes), s_queue its per-wave queues.  Starts
// with a barrier, so a workgroup may call it for tile after tile.
// LISTS: walk the candidate lists of the tile's table (SampleTable::cl_*).
// DEV_TABLE: read the table from a.level_tables (the single-plane kernel)
// rather than st_all (the batch kernel's plane record).
template <int BLOCK, bool LISTS, bool DEV_TABLE>
__device__ __forceinline__ void splat_tile(const PixelArgs &a, const PixGeom &geo, const SampleTable &st_all,
                                           const uint32_t *bin_start, const float4 *grains, int tile,
                                           unsigned char *lds, uint32_t (*s_queue)[kQueueCap],
                                           float4 (*s_grec)[64])
{
    constexpr int NW = BLOCK / 64;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    __syncthreads();
    if (tile >= a.n_tiles) return;
    if (a.grain_total && *a.grain_total > a.grain_cap) return;
    if (tile_cancelled(a.cancel, a.status)) return;
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
    t.I0 = geo.Ix0;  // grain cells are packed relative to the extended grid
    t.J0 = geo.Jy0;
    const float rm = a.rm, delta = a.delta, rd = a.rdelta, iz = a.inv_zoom, z = a.zoom;
    const int fast = a.div_fast;
    // adaptive sampling: the tile's level table holds its first N samples
    // The single-plane kernel (DEV_TABLE) reads its table from a.level_tables:
    // the tile's level's table, or the one table of a fixed-N render.  It is
    // read through the constant address space (the tile is uniform, so the
    // level and the table's fields become scalar loads, as kernel arguments)
    // and never selected at run time against st_all, which would make the
    // compiler keep a copy of the walk for each.  The batch kernel passes its
    // plane's table as st_all.
    const SampleTable *lt = &st_all;
    if (DEV_TABLE) {
        const int lvl = a.budget ? ((const __attribute__((address_space(4))) uint8_t *)a.tile_level)[tile] : 0;
        lt = (const SampleTable *)((const __attribute__((address_space(4))) SampleTable *)a.level_tables + lvl);
    }
    const SampleTable st = *lt;  // a copy: the walk's LDS atomics cannot make it reload
    const int N = st.n, words = a.words, Q = st.qbuckets;
    const TileRegion reg = tile_region(a, geo, t.X0, t.Y0, t.X1, t.Y1);
    const int bx0 = reg.bx0, bx1 = reg.bx1, by0 = reg.by0, by1 = reg.by1;
    const int rci0 = reg.ci0, rci1 = reg.ci1, rcj0 = reg.cj0, rcj1 = reg.cj1;

    // ---- LDS: (o, word, bit)[N] | qidx[Q + 1] | bitset ----
    // Sample k's bit of tile pixel i is bit k % 32 of bitset word i * words +
    // k / 32.  Per sorted position one 16-byte record holds the sample's
    // offset in input pixels, that word's LDS byte offset for pixel 0 and the
    // bit's mask: a test reads it with one load, and a hit's address is one
    // multiply-add.  (Records of 24 bytes that also held the scaled offset,
    // read with two loads, made the splat 9% slower.)
    const int tile_words = a.tile_w * a.tile_h * words;
    const uint32_t w4 = (uint32_t)words * 4u;
    const uint32_t row_b = (uint32_t)t.TW * w4;  // bitset bytes of a tile row
    // per sorted position one 16-byte record: the offset in input pixels,
    // the LDS address of the bitset word of its sample for pixel 0 and the
    // bit's mask; the scaled offset X = O * zoom is recomputed (the table's
    // own f32 product).  First in LDS: at a fixed address, a record is read
    // at its byte offset alone (the candidate lists hold p * 16).
    size_t off = 0;
    float4 *srec = reinterpret_cast<float4 *>(lds);
    off += align16((size_t)N * 16);
    uint16_t *sq = reinterpret_cast<uint16_t *>(lds + off);
    off += align16((size_t)(Q + 1) * 2);
    uint32_t *bits = reinterpret_cast<uint32_t *>(lds + off);
    off += align16((size_t)tile_words * 4);
    if (off > (size_t)a.lds_bytes) {
        if (tid == 0) atomicOr(a.status, kStatusRegion);
        return;
    }
    for (int w = tid; w < tile_words; w += BLOCK) bits[w] = 0u;
    for (int k = tid; k < N; k += BLOCK) {
        const float2 O = st.o[k];
        const uint32_t ki = st.k[k];
        srec[k] = make_float4(O.x, O.y, __uint_as_float(lds_addr(bits + (ki >> 5))), __uint_as_float(1u << (ki & 31u)));
    }
    for (int q = tid; q <= Q; q += BLOCK) sq[q] = st.qidx[q];
    // (X, O) of sorted position p: the scaled offset and the offset in
    // input pixels
    auto XO_at = [&](int p) {
        const float4 r = srec[p];
        return make_float4(r.x * z, r.y * z, r.x, r.y);
    };
    // the record at byte offset pb (= 16 p)
    auto rec_at = [&](uint32_t pb) {
        return *reinterpret_cast<const float4 *>(lds + pb);
    };
    // the (word byte offset, bit mask) pair of sorted position p
    auto word_bit = [&](int p) {
        const float4 r = srec[p];
        return make_uint2(__float_as_uint(r.z), __float_as_uint(r.w));
    };
    const float4 ext = *st.ext;
    const RowBuckets rbk = row_buckets(ext);
    // the tile's 64-grain chunks go to its waves on demand (the first NW
    // statically), so that a wave that drew sparse chunks takes more
    __shared__ int s_next_chunk;
    if (tid == 0) s_next_chunk = NW;
    __syncthreads();

    const float slack = 4e-3f + 2e-6f * (float)max(a.out_w, a.out_h);
    const float eps = st.eps, safe_r = st.safe_r, Qf = (float)Q;
    // A pixel's point is (x + 0.5) * iz - O.  For integral x, x + 0.5 is exact,
    // so fma(x, iz, iz / 2) rounds the same exact product once: the same f32.
    const float hiz = 0.5f * iz;
    uint32_t *queue = s_queue[wid];
    int qcount = 0;

    // Per-lane state of the wave's current 64-grain chunk: lane l holds grain
    // l's centre, signed radius (negative: the cell window must be checked)
    // and packed cell.
    float gcx = 0.0f, gcy = 0.0f, gr = 0.0f;
    uint32_t gcell = 0;

    // A queue entry is one (grain, sample, pixel) test: grain lane (6 bits) |
    // sorted sample position << 6 (11 bits) | tile column << 17 (7 bits) |
    // tile row << 24 (8 bits).
    // Tests entries [0, n) against the reference's disk test and cell window
    // and moves entries [64, qcount) to the front.  Called by the whole wave.
    auto drain = [&](int n) {
        uint32_t e = 0;
        if (lane < n) e = queue[lane];
        if (lane + 64 < qcount) queue[lane] = queue[lane + 64];
        const int l = (int)(e & 63u);
        const float cx = __shfl(gcx, l, 64), cy = __shfl(gcy, l, 64), rs = __shfl(gr, l, 64);
        const uint32_t cell = (uint32_t)__shfl((int)gcell, l, 64);
        if (lane >= n) return;
        const int p = (int)((e >> 6) & 0x7FFu), tx = (int)((e >> 17) & 0x7Fu), ty = (int)(e >> 24);
        const float4 XO = XO_at(p);
        const float2 O = make_float2(XO.z, XO.w);
        const uint2 wm = word_bit(p);
        const float xg = __builtin_fmaf((float)(t.X0 + tx), iz, hiz) - O.x;
        const float yg = __builtin_fmaf((float)(t.Y0 + ty), iz, hiz) - O.y;
        const float dx = xg - cx, dy = yg - cy;
        const float r = fabsf(rs);
        if (dx * dx + dy * dy <= r * r) {
            const int gi = t.I0 + (int)(cell & 0xFFFFu), gj = t.J0 + (int)(cell >> 16);
            if (rs > 0.0f ||
                (floor_div(xg - rm, delta, rd, fast) <= gi && gi <= floor_div(xg + rm, delta, rd, fast) &&
                 floor_div(yg - rm, delta, rd, fast) <= gj && gj <= floor_div(yg + rm, delta, rd, fast)))
                lds_or(mad_u24((uint32_t)ty, row_b, mad_u24((uint32_t)tx, w4, wm.x)), wm.y);
        }
    };

    // Appends the entries of the lanes with cand set (at most 63 + 64 are held).
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

    // The direct test of a narrow grain whose disk test implies its cell
    // window (rs > 0) at the sample with record r (offset O, bitset word
    // address and bit).  A pixel whose point passes the disk test lies
    // within rz of the sample's target (bx + X.x, by + X.y), X = O * zoom, in
    // each axis (the candidate box is a superset); with rz < 0.5 that is
    // only the pixel nearest the target.  So the drain's test (same f32
    // expression, rs2 = r * r) of that one pixel, when it lies in the tile,
    // sets exactly the bits the queue would have set.  The target is taken
    // with one rounding, fma(O, zoom, bx): it moves by far less than the
    // slack in rz, so the nearest pixel is the same whenever one can pass.
    // The grain's values may differ between lanes.
    const int tw = t.X1 - t.X0, th = t.Y1 - t.Y0;
    auto test = [&](bool ok, float4 r, float gbx, float gby, float gcx_, float gcy_, float rs2) {
        const float px = rintf(__builtin_fmaf(r.x, z, gbx)), py = rintf(__builtin_fmaf(r.y, z, gby));
        const int tx = (int)px - t.X0, ty = (int)py - t.Y0;
        if (ok & ((unsigned)tx < (unsigned)tw) & ((unsigned)ty < (unsigned)th)) {
            const float xg = __builtin_fmaf(px, iz, hiz) - r.x;
            const float yg = __builtin_fmaf(py, iz, hiz) - r.y;
            const float dx = xg - gcx_, dy = yg - gcy_;
            if (dx * dx + dy * dy <= rs2)
                // tile coordinates and the bitset's bytes are far below 2^24:
                // two 24-bit multiply-adds (full rate; the compiler otherwise
                // forms the pixel index with a quarter-rate 64-bit mad)
                lds_or(mad_u24((uint32_t)ty, row_b, mad_u24((uint32_t)tx, w4, __float_as_uint(r.z))),
                       __float_as_uint(r.w));
        }
    };

    // ---- each wave walks the region's bin rows in 64-grain chunks ----
    int row = by0, cbase = 0, nch = 0;
    uint32_t rs_ = 0, re_ = 0;
    auto row_load = [&](int by) {
        rs_ = bin_start[by * geo.nbx + bx0];
        re_ = bin_start[by * geo.nbx + bx1 + 1];
        nch = re_ > rs_ ? (int)((re_ - rs_ + 63u) / 64u) : 0;
    };
    if (bx0 <= bx1 && row <= by1) row_load(row);
    else row = by1 + 1;
    int c = wid;
    while (row <= by1) {
        if (c >= cbase + nch) {
            cbase += nch;
            if (++row <= by1) row_load(row);
            continue;
        }
        const uint32_t q = rs_ + (uint32_t)(c - cbase) * 64u + (uint32_t)lane;
        {
            int v = 0;
            if (lane == 0) v = atomicAdd(&s_next_chunk, 1);
            c = __builtin_amdgcn_readfirstlane(v);
        }
        float4 G = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const bool in = q < re_;
        // grain records stream through once per tile (the halo re-reads of
        // other tiles come much later): non-temporal, so that they do not
        // push the candidate lists out of L2
        if (in) {
            typedef float f4 __attribute__((ext_vector_type(4)));
            const f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(grains + q));
            G = make_float4(v.x, v.y, v.z, v.w);
        }
        const uint32_t pk = __float_as_uint(G.w);
        const int ci = (int)(pk & 0xFFFFu), cj = (int)(pk >> 16);
        const float rad = G.z;
        const float bx = G.x * z - 0.5f, by = G.y * z - 0.5f, rz = rad * z + slack;
        // a grain outside the region enters no window of the tile; one whose
        // boxes miss the tile for every sample cannot cover any of its points
        bool keep = in && rad > 0.0f && ci >= rci0 && ci <= rci1 && cj >= rcj0 && cj <= rcj1;
        keep = keep && floorf((bx + ext.y) + rz) >= t.fx0 && ceilf((bx + ext.x) - rz) <= t.fx1 &&
               floorf((by + ext.w) + rz) >= t.fy0 && ceilf((by + ext.z) - rz) <= t.fy1;
        // The box has a column iff frac(X) lies in [A - 2rz, A] (mod 1) with
        // A = frac(rz - bx); widened by eps and rounded out to whole buckets.
        int p0 = 0, len = 0;
        if (keep) {
            const float nc = rz - bx;
            const float A = nc - floorf(nc);
            const float lo = (A - (rz + rz)) - eps, hi = A + eps;
            int lq = (int)floorf(lo * Qf), hq = (int)floorf(hi * Qf) + 1;
            if (!(lo * Qf > -2.0f * Qf) || !(hi * Qf < 2.0f * Qf) || hq - lq >= Q) {
                len = N;
            } else {
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
        // a narrow grain that takes the direct test and has a radius class:
        // its cell's candidate list replaces the run
        bool listed = false;
        if (LISTS && keep && st.cl_cell && rad <= safe_r && rz + rz < 1.0f) {
            const int k = (int)(rz * a.cl_inv_dz);
            if (k < a.cl_classes) {
                const float ncx = -bx, ncy = -by;
                const float fx = ncx - floorf(ncx), fy = ncy - floorf(ncy);
                const float qf = (float)a.cl_q;
                const int qa = min((int)(fx * qf), a.cl_q - 1), qb = min((int)(fy * qf), a.cl_q - 1);
                const size_t cix = (size_t)k * (a.cl_q * a.cl_q) + qb * a.cl_q + qa;
                const uint2 se = st.cl_cell[cix];
                // Only the list's entries whose target row can be a tile
                // row: rint(by + X.y) in [Y0, Y1 - 1] needs X.y in
                // [Y0 - 0.5 - by, Y1 - 0.5 - by].  Widened by 2^-4 for the
                // rounding of the target and of the bounds, the row buckets
                // of that interval are one run of the list's entries.
                const int b0 = row_bucket(rbk, (t.fy0 - 0.5f - by) - 0.0625f);
                const int b1 = row_bucket(rbk, (t.fy1 + 0.5f - by) + 0.0625f);
                const uint16_t *yx = st.cl_yidx + cix * kRowBuckets;
                const int e0 = b0 > 0 ? (int)yx[b0 - 1] : 0;
                p0 = (int)se.x + e0;
                len = (int)yx[b1] - e0;
                listed = true;
            }
        }
        gcx = G.x;
        gcy = G.y;
        gr = rad <= safe_r ? rad : -rad;
        gcell = pk;
        // the listed grains (all narrow, all with rs > 0) take the
        // concatenated walk below; the others the per-grain walk after it
        unsigned long long m = __ballot(keep && !listed);
        unsigned long long lrem = LISTS ? __ballot(listed && len > 0) : 0ull;
        if (LISTS && lrem) {
            // The listed grains' candidate lists, concatenated in lane order,
            // are walked 64 entries a step: stream entry e belongs to the
            // last listed grain whose first entry excl is at most e, and is
            // its list entry cl_list[p0 + e - excl].
            const uint32_t fl = listed ? (uint32_t)len : 0u;
            const uint32_t incl = wave_inclusive_sum(fl);
            const int total = lane_i((int)incl, 63);
            const int excl = (int)(incl - fl);
            // Entry e of the stream is list entry dl + e of its owner grain,
            // dl = p0 - excl >= -excl > -total.  A step from s loads entry
            // s + min(lane, total - 1 - s), and the steps of a batch that
            // start past the stream have s < total + 64 (kWalkStage - 1), so
            // biased by bias = total + 64 kWalkStage, the lane's part
            // dlb + min(lane, total - 1 - s) is positive: a valid unsigned
            // 32-bit offset from the batch's base.
            const int bias = total + 64 * kWalkStage;
            const int dlb = p0 - excl + bias;
            const float rs2 = rad * rad;
            // the chunk's grains (target origin, centre) in this wave's LDS
            // records: a step reads its owner's with one load instead of
            // broadcasting four values, and its r^2 with one lane read (the
            // wave's LDS accesses run in order, so no barrier is needed)
            s_grec[wid][lane] = make_float4(bx, by, G.x, G.y);
            const unsigned char *grec = reinterpret_cast<const unsigned char *>(&s_grec[wid][0]);
            int carry = 0;  // 4 x the last grain that starts before the step
            // 4 x the owner of each lane's entry of the step from s0 (a lane
            // read's byte address)
            auto owner_of = [&](int s0) {
                const int e = s0 + lane;
                int owner = carry;
                while (lrem) {
                    const int g = __ffsll((long long)lrem) - 1;
                    const int ex = lane_i(excl, g);
                    if (ex >= s0 + 64) break;
                    owner = e >= ex ? (g << 2) : owner;
                    carry = g << 2;
                    lrem &= lrem - 1ull;
                }
                return owner;
            };
            // Lanes past the stream read its last entry (the stream's last
            // pair again: its bit is ORed twice, which changes nothing).  The
            // load takes a base in SGPRs (the batch's first entry, less the
            // bias), a 32-bit byte offset in one VGPR and the step's 128
            // bytes as its immediate offset: no 64-bit address arithmetic.
            auto entry = [&](const unsigned char *lbb, int i, int s0, int own4) {
                const uint32_t pos = (uint32_t)from_lane_i(dlb, own4) + (uint32_t)min(lane, total - 1 - s0);
                return (uint32_t)*reinterpret_cast<const uint16_t *>(lbb + 128 * i + (pos << 1));
            };
            // kWalkStage steps at a time: their owners first, then all their
            // list loads in flight together (steps past the stream read its
            // last entry and test nothing), then their tests: one global
            // latency per batch of steps instead of one per step
            for (int sb = 0; sb < total; sb += 64 * kWalkStage) {
                const unsigned char *lbb =
                    reinterpret_cast<const unsigned char *>(st.cl_list) + 2 * ((long long)sb - bias);
                int own[kWalkStage];
                uint32_t pn[kWalkStage];
#pragma unroll
                for (int i = 0; i < kWalkStage; ++i) own[i] = owner_of(sb + 64 * i);
#pragma unroll
                for (int i = 0; i < kWalkStage; ++i) pn[i] = entry(lbb, i, sb + 64 * i, own[i]);
#pragma unroll
                for (int i = 0; i < kWalkStage; ++i) {
                    const int s0 = sb + 64 * i;
                    if (s0 >= total) break;
                    // the owner grain's target origin and centre from its
                    // record, its r^2 from its lane; the entry is the
                    // sample record's byte offset
                    const float4 gr = *reinterpret_cast<const float4 *>(grec + (own[i] << 2));
                    const float ors2 = from_lane_f(rs2, own[i]);
                    const float4 r = rec_at(pn[i]);
                    // both reads issue before the test's branches (as in the
                    // per-grain walk)
                    asm volatile("" ::"v"(r.x), "v"(r.y), "v"(r.z), "v"(r.w));
                    test(true, r, gr.x, gr.y, gr.z, gr.w, ors2);
                }
            }
        }
        while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1ull;
            const float ubx = lane_f(bx, l), uby = lane_f(by, l), urz = lane_f(rz, l);
            const int up0 = lane_i(p0, l), ulen = lane_i(len, l);
            // A box narrower than one pixel (2 rz < 1) holds at most one
            // pixel column and one row: decided once per grain.
            const bool narrow = urz + urz < 1.0f;
            // Such a grain with no cell window to check (rs > 0) takes the
            // reference's disk test of its one pixel in the walk itself.
            const float urs = lane_f(gr, l);
            const bool direct = narrow && urs > 0.0f;
            // One sample of the run per lane (i: index in the run, p: its
            // sorted position, X: its scaled offset).
            auto step = [&](int i, int p, float2 X) {
                // the candidate box (splat_box): a point of the grain's disk
                // is a pixel in it; the run makes its column range non-empty
                // up to the widening, and the drain's exact test settles the
                // rest, so only the rows and the tile are checked here
                const float cxk = ubx + X.x, cyk = uby + X.y;
                const float xl = ceilf(cxk - urz), xh = floorf(cxk + urz);
                const float yl = ceilf(cyk - urz), yh = floorf(cyk + urz);
                const bool cand = (i < ulen) & (yl <= yh) & (xl <= t.fx1) & (xh >= t.fx0) & (yl <= t.fy1) &
                                  (yh >= t.fy0);
                const uint32_t base = (uint32_t)l | ((uint32_t)p << 6);
                if (narrow || !__any(cand & ((xh > xl) | (yh > yl)))) {
                    const int tx = (int)(xl - t.fx0), ty = (int)(yl - t.fy0);
                    push(cand, base | ((uint32_t)tx << 17) | ((uint32_t)ty << 24));
                } else {
                    // boxes of more than one pixel: one entry per pixel of
                    // the box clipped to the tile
                    const float cxl = fmaxf(xl, t.fx0), cxh = fminf(xh, t.fx1);
                    const float cyl = fmaxf(yl, t.fy0), cyh = fminf(yh, t.fy1);
                    const bool c2 = cand & (cxl <= cxh);
                    const int bw = c2 ? (int)(cxh - cxl) + 1 : 0, bh = c2 ? (int)(cyh - cyl) + 1 : 0;
                    const int nb = bw * bh;
                    const int tx = c2 ? (int)(cxl - t.fx0) : 0, ty = c2 ? (int)(cyl - t.fy0) : 0;
                    for (int j = 0; __any(j < nb); ++j) {
                        const bool c = j < nb;
                        const int jy = c ? j / bw : 0, jx = c ? j - jy * bw : 0;
                        push(c, base | ((uint32_t)(tx + jx) << 17) | ((uint32_t)(ty + jy) << 24));
                    }
                }
            };
            // cyclic position: pr - N when pr >= N (as unsigned, pr - N wraps
            // above pr otherwise); lanes past the run read the last sample
            // and queue nothing
            auto position = [&](int pr) {
                return (int)min(min((uint32_t)pr, (uint32_t)(pr - N)), (uint32_t)(N - 1));
            };
            // two 64-sample halves per iteration: both halves' LDS reads are
            // issued before either is tested
            if (direct) {
                const float ucx = lane_f(gcx, l), ucy = lane_f(gcy, l), urs2 = urs * urs;
                for (int i0 = 0; i0 < ulen; i0 += 128) {
                    const int ia = i0 + lane, ib = ia + 64;
                    const int pa = position(up0 + ia), pb = position(up0 + ib);
                    const float4 ra = srec[pa], rb = srec[pb];
                    // all the reads issue together, ahead of the tests (an
                    // empty asm that takes the values as operands keeps the
                    // compiler from moving the record reads into the
                    // branches after the dependent ones)
                    asm volatile("" ::"v"(ra.x), "v"(ra.y), "v"(ra.z), "v"(ra.w), "v"(rb.x), "v"(rb.y),
                                 "v"(rb.z), "v"(rb.w));
                    test(ia < ulen, ra, ubx, uby, ucx, ucy, urs2);
                    if (i0 + 64 < ulen) test(ib < ulen, rb, ubx, uby, ucx, ucy, urs2);
                }
                continue;
            }
            for (int i0 = 0; i0 < ulen; i0 += 128) {
                const int ia = i0 + lane, ib = ia + 64;
                const int pa = position(up0 + ia), pb = position(up0 + ib);
                const float4 Xa = XO_at(pa), Xb = XO_at(pb);
                step(ia, pa, make_float2(Xa.x, Xa.y));
                if (i0 + 64 < ulen) step(ib, pb, make_float2(Xb.x, Xb.y));
            }
        }
        // the queue refers to this chunk's lanes: empty it before the next
        if (qcount > 0) {
            drain(qcount);
            qcount = 0;
        }
    }
    __syncthreads();

    // ---- popcount epilogue: one store per output pixel ----
    for (int p = tid; p < tw * th; p += BLOCK) {
        const int ox = p % tw, oy = p / tw;
        const uint32_t *pw = bits + ((size_t)oy * a.tile_w + ox) * words;
        uint32_t count = 0;
        const size_t po = (size_t)(t.Y0 + oy) * a.out_w + (t.X0 + ox);
        if (a.budget) {
            // the pixel's first n samples, as a render with N = n counts them
            const int n = min(max((int)a.budget[po], 1), N);
            for (int w = 0; w < (n >> 5); ++w) count += __popc(pw[w]);
            if (n & 31) count += __popc(pw[n >> 5] & ((1u << (n & 31)) - 1u));
            a.out[po] = (float)count * (1.0f / (float)n);
        } else {
            for (int w = 0; w < words; ++w) count += __popc(pw[w]);
            a.out[po] = (float)count * a.inv_samples;
        }
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 6) void splat_hbm_kernel(PixelArgs a, PixGeom geo, SampleTable st,
                                                          const uint32_t *bin_start, const float4 *grains)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t s_queue[BLOCK / 64][kQueueCap];
    __shared__ float4 s_grec[BLOCK / 64][64];
    const int tile = a.tile_list ? a.tile_list[blockIdx.x] : (int)blockIdx.x;
    splat_tile<BLOCK, true, true>(a, geo, st, bin_start, grains, tile, lds, s_queue, s_grec);
}

// Persistent batch kernel: every workgroup takes (plane, tile) items from the
// batch's launch order through an atomic counter until the items run out
// (each workgroup exits after reading an index past the end).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 6) void splat_batch_kernel(const BatchPlane *planes, const int2 *items, int n_items,
                                                            unsigned int *counter)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t s_queue[BLOCK / 64][kQueueCap];
    __shared__ float4 s_grec[BLOCK / 64][64];
    __shared__ int s_item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        // the item is uniform: say so, so that the plane's fields are scalar loads
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= n_items) break;
        const int2 pt = items[it];
        const int pl = __builtin_amdgcn_readfirstlane(pt.x), tile = __builtin_amdgcn_readfirstlane(pt.y);
        // No thread writes the plane table while the kernel runs: reached
        // through the constant address space, its fields at this uniform
        // address become scalar loads into SGPRs, as kernel arguments are
        // (generic loads kept them in VGPRs: 91 VGPRs against 68).
        const auto *cp = (const __attribute__((address_space(4))) BatchPlane *)planes + pl;
        const BatchPlane &p = *(const BatchPlane *)cp;
        splat_tile<BLOCK, true, false>(p.a, p.geo, p.st, p.bin_start, p.grains, tile, lds, s_queue, s_grec);
    }
}

__global__ void store_frame_planes_kernel(FramePlaneRecords r, int n, BatchPlane *dst)
{
    if (blockIdx.x == 0 && (int)threadIdx.x < n) dst[threadIdx.x] = r.p[threadIdx.x];
}

hipError_t launch_store_frame_planes(const FramePlaneRecords &r, int n, BatchPlane *dst, hipStream_t s)
{
    if (n < 1 || n > kFramePlanes) return hipErrorInvalidValue;
    hipLaunchKernelGGL(store_frame_planes_kernel, dim3(1), dim3(64), 0, s, r, n, dst);
    return hipGetLastError();
}

int splat_batch_blocks_per_cu(int block, int lds_bytes)
{
    int nb = 0;
    const void *f = block == 512 ? reinterpret_cast<const void *>(&splat_batch_kernel<512>)
                                 : reinterpret_cast<const void *>(&splat_batch_kernel<256>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, block, (size_t)lds_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return nb > 0 ? nb : 1;
}

hipError_t launch_splat_batch(const BatchPlane *planes, const int2 *items, int n_items, unsigned int *counter,
                              int grid, int block, int lds_bytes, hipStream_t s)
{
    if (block == 512)
        hipLaunchKernelGGL((splat_batch_kernel<512>), dim3(grid), dim3(512), (size_t)lds_bytes, s, planes, items,
                           n_items, counter);
    else
        hipLaunchKernelGGL((splat_batch_kernel<256>), dim3(grid), dim3(256), (size_t)lds_bytes, s, planes, items,
                           n_items, counter);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void lambda_check_kernel(const float *lambda, long long n, int *status)
{
    bool bad = false;
    for (long long p = (long long)blockIdx.x * kBlock + threadIdx.x; p < n; p += (long long)gridDim.x * kBlock)
        bad = bad || lambda_invalid(lambda[p]);
    if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(status, kStatusLambda);
}

hipError_t launch_lambda_check(const float *lambda, long long n, int *status, hipStream_t s)
{
    const long long blocks = std::min<long long>((n + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(lambda_check_kernel, dim3((unsigned)std::max<long long>(blocks, 1)), dim3(kBlock), 0, s,
                       lambda, n, status);
    return hipGetLastError();
}

hipError_t launch_threshold(const PixelArgs &a, double *T, hipStream_t s)
{
    const long long n = (long long)a.in_w * a.in_h;
    hipLaunchKernelGGL(threshold_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       a.lambda, n, a.delta, T, a.status);
    return hipGetLastError();
}

// chunk bins a wave
static dim3 bin_grid(const PixGeom &g, int chunk)
{
    const long long waves = ((long long)g.nbx * g.nby + chunk - 1) / chunk;
    return dim3((unsigned)((waves + kBlock / 64 - 1) / (kBlock / 64)));
}

hipError_t launch_bin_count(const PixelArgs &a, const PixGeom &g, const double *T, uint32_t *bin_count,
                            unsigned long long *bin_mask, hipStream_t s)
{
    FramePlanes fp{};
    fp.T[0] = T;
    fp.cnt[0] = bin_count;
    fp.mask[0] = bin_mask;
    hipLaunchKernelGGL(bin_count_planes_kernel<1>, bin_grid(g, 1), dim3(kBlock), 0, s, a, g, fp);
    return hipGetLastError();
}

hipError_t launch_bin_count_planes(const PixelArgs &a, const PixGeom &g, const FramePlanes &fp, int n_planes,
                                   hipStream_t s)
{
    switch (n_planes) {
    case 1: hipLaunchKernelGGL(bin_count_planes_kernel<1>, bin_grid(g, 1), dim3(kBlock), 0, s, a, g, fp); break;
    case 2: hipLaunchKernelGGL(bin_count_planes_kernel<2>, bin_grid(g, 1), dim3(kBlock), 0, s, a, g, fp); break;
    case 3: hipLaunchKernelGGL(bin_count_planes_kernel<3>, bin_grid(g, 1), dim3(kBlock), 0, s, a, g, fp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t scan_bins(void *temp, size_t *temp_bytes, const uint32_t *in, uint32_t *out, int n, hipStream_t s)
{
    return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, in, out, n, s);
}

hipError_t launch_bin_grains(const PixelArgs &a, const PixGeom &g, int dist_kind, const double *T,
                             const uint32_t *bin_start, const unsigned long long *bin_mask, float4 *grains,
                             hipStream_t s)
{
    if (dist_kind == 0)
        hipLaunchKernelGGL((bin_grains_kernel<0>), bin_grid(g, kDrawChunk), dim3(kBlock), 0, s, a, g, T, bin_start, bin_mask,
                           grains);
    else
        hipLaunchKernelGGL((bin_grains_kernel<1>), bin_grid(g, kDrawChunk), dim3(kBlock), 0, s, a, g, T, bin_start, bin_mask,
                           grains);
    return hipGetLastError();
}

static dim3 wave_grid(long long waves) { return dim3((unsigned)((waves + kBlock / 64 - 1) / (kBlock / 64))); }

hipError_t launch_batch_threshold(const BatchGen &bg, hipStream_t s)
{
    if (bg.thr_blocks <= 0 || bg.thr_blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_threshold_kernel, dim3((unsigned)bg.thr_blocks), dim3(kBlock), 0, s, bg);
    return hipGetLastError();
}

hipError_t launch_batch_bin_grains(const BatchGen &bg, hipStream_t s)
{
    if (bg.kind_chunks[0] > 0)
        hipLaunchKernelGGL(batch_bin_grains_kernel<0>, wave_grid(bg.kind_chunks[0]), dim3(kBlock), 0, s, bg);
    if (bg.kind_chunks[1] > 0)
        hipLaunchKernelGGL(batch_bin_grains_kernel<1>, wave_grid(bg.kind_chunks[1]), dim3(kBlock), 0, s, bg);
    return hipGetLastError();
}

hipError_t launch_batch_sample_tables(const BatchGen &bg, hipStream_t s)
{
    if (bg.n <= 0 || bg.n > 65535 || bg.sort_blocks <= 0 || bg.list_blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_sample_sort_kernel, dim3((unsigned)bg.sort_blocks, (unsigned)bg.n), dim3(kBlock), 0, s,
                       bg);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_sample_index_kernel, dim3((unsigned)bg.n), dim3(kBlock), 0, s, bg);
    e = hipGetLastError();
    if (e != hipSuccess || bg.list_blocks == 0) return e;
    hipLaunchKernelGGL(batch_splat_lists_kernel, dim3((unsigned)bg.list_blocks), dim3(256), 0, s, bg);
    return hipGetLastError();
}

hipError_t launch_splat_hbm(const PixelArgs &a, const PixGeom &g, const SampleTable &st, const uint32_t *bin_start,
                            const float4 *grains, int block, hipStream_t s)
{
    const dim3 grid(a.n_tiles);
    const size_t lds = (size_t)a.lds_bytes;
    if (block == 512)
        hipLaunchKernelGGL((splat_hbm_kernel<512>), grid, dim3(512), lds, s, a, g, st, bin_start, grains);
    else
        hipLaunchKernelGGL((splat_hbm_kernel<256>), grid, dim3(256), lds, s, a, g, st, bin_start, grains);
    return hipGetLastError();
}

}  // namespace fg
