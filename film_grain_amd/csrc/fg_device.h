// fg_device.h -- device helpers shared by the film-grain kernels.
#ifndef FG_DEVICE_H
#define FG_DEVICE_H

#include "fg_kernels.h"
#include "fg_rng.h"

namespace fg {

static __constant__ double d_zig_x[257] = {FG_ZIG_NORM_X_VALUES};
static __constant__ double d_zig_f[257] = {FG_ZIG_NORM_F_VALUES};

// floor(RN(a / d)) cast like Rust's `as i32`.  With fast set, the quotient
// comes from q0 = a*rd corrected by one FMA residual step, which the host has
// verified equals the IEEE quotient for this d over a whole binade; outside
// [2^-60, 2^60] (where intermediates could leave the normal range) the IEEE
// division is used.
__device__ __forceinline__ int floor_div(float a, float d, float rd, int fast)
{
    float q;
    const float aa = fabsf(a);
    if (fast && aa >= 8.67361737988404e-19f && aa <= 1.152921504606847e18f) {
        const float q0 = a * rd;
        const float r = fmaf(-q0, d, a);
        q = fmaf(r, rd, q0);
    } else {
        q = __fdiv_rn(a, d);
    }
    return sat_i32(floorf(q));
}

// Workgroup exclusive scan of one u32 per thread (kBlock threads); returns the
// exclusive prefix and the workgroup total.  Contains two barriers.
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

// Poisson threshold of one lambda value for the pixel-wise cells
// (src/pixelwise.rs:69-82): 2.0 when the cell is skipped (lambda <= 0 or
// lambda*delta*delta <= 0), exp(-expected) for the Knuth branch (expected < 12),
// and -expected for the rejection branch.  Values outside the accepted range
// (NaN, above kMaxLambda) are skipped too; lambda_invalid flags them.
__device__ __forceinline__ bool lambda_invalid(float lam) { return !(lam <= kMaxLambda); }
__device__ __forceinline__ double cell_threshold(float lam, float delta)
{
    if (!(lam > 0.0f) || !(lam <= kMaxLambda)) return 2.0;
    const float expected = lam * delta * delta;
    if (!(expected > 0.0f)) return 2.0;
    if (expected < 12.0f) return exp_neg_glibc(expected);  // glibc's exp(-expected), exactly
    return -(double)expected;
}

// Poisson count of a cell given its hash and threshold, leaving rng positioned
// after the count draws (as Poisson::sample leaves it).  The common empty cell
// costs only the two seed words the first Xoshiro output depends on.
__device__ __forceinline__ uint32_t cell_count(uint64_t h, double e, Xoshiro &rng)
{
    if (e > 1.5) return 0u;
    if (e >= 0.0) {
        const uint64_t s0 = splitmix64(h), s3 = splitmix64(h + 3 * kGolden);
        const double u1 = standard_f64(rotl64(s0 + s3, 23) + s0);
        if (!(u1 > e)) return 0u;  // Knuth: p = u1 <= exp(-lambda) stops at once
        rng.seed(h);
        (void)rng.next();
        double p = u1;
        uint32_t n = 1;
        while (p > e) {
            p = p * standard_f64(rng.next());
            ++n;
        }
        return n - 1u;
    }
    rng.seed(h);
    const double r = poisson_rejection(rng, -e);
    if (!(r > 0.0)) return 0u;
    if (r >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)r;
}

// ---------------------------------------------------------------------------
// splat drain, shared by the tile and HBM splat kernels
// ---------------------------------------------------------------------------
struct SplatTile {
    int X0, Y0, X1, Y1;  // output rect
    int I0, J0;          // region origin (cell indices)
    int TW;              // bitset row pitch (tile_w)
    float fx0, fx1, fy0, fy1;  // X0, X1-1, Y0, Y1-1 as floats
};

// Candidate box of grain record B = (cx*zoom - 0.5, cy*zoom - 0.5, r*zoom +
// slack) for sample k with scaled offset (oxz, oyz): the output pixels whose
// sample point can lie in the grain's disk.  The phase that queues candidates
// and the drain that tests them evaluate this with the same operations, so
// they agree on every box.
struct SplatBox {
    float xl, xh, yl, yh;
};
__device__ __forceinline__ SplatBox splat_box(const float4 &B, float oxz, float oyz)
{
    const float cx = B.x + oxz, cy = B.y + oyz;
    SplatBox b;
    b.xl = ceilf(cx - B.z);
    b.xh = floorf(cx + B.z);
    b.yl = ceilf(cy - B.z);
    b.yh = floorf(cy + B.z);
    return b;
}
// The box is non-empty and meets the tile.
__device__ __forceinline__ bool splat_box_hits(const SplatBox &b, const SplatTile &t)
{
    return b.xl <= b.xh && b.xl <= t.fx1 && b.xh >= t.fx0 && b.yl <= b.yh && b.yl <= t.fy1 &&
           b.yh >= t.fy0;
}

// Drains the first n (<= 64) entries (grain | k << 16) of this wave's queue
// through the exact reference test and moves entries [64, count) to the front.
__device__ __forceinline__ void splat_drain(const PixelArgs &a, const SplatTile &t, uint32_t *queue,
                                            int n, int count, const float4 *gA, const float4 *gB,
                                            const float2 *offs, uint32_t *bits)
{
    const int lane = threadIdx.x & 63;
    uint32_t e = 0;
    if (lane < n) e = queue[lane];
    if (lane + 64 < count) queue[lane] = queue[lane + 64];
    if (lane >= n) return;
    const uint32_t g = e & 0xFFFFu;
    const int k = (int)(e >> 16);
    const float2 o = offs[k];
    const SplatBox b = splat_box(gB[g], o.x * a.zoom, o.y * a.zoom);
    const int xlo = (int)fmaxf(b.xl, t.fx0), xhi = (int)fminf(b.xh, t.fx1);
    const int ylo = (int)fmaxf(b.yl, t.fy0), yhi = (int)fminf(b.yh, t.fy1);
    const float4 G = gA[g];
    const uint32_t cell = __float_as_uint(G.w);
    const int gi = t.I0 + (int)(cell & 0xFFFFu), gj = t.J0 + (int)(cell >> 16);
    const float rm = a.rm, delta = a.delta, rd = a.rdelta, iz = a.inv_zoom;
    const int fast = a.div_fast;
    for (int y = ylo; y <= yhi; ++y) {
        const float yg = ((float)y + 0.5f) * iz - o.y;
        const float dy = yg - G.y;
        const float dy2 = dy * dy;
        for (int x = xlo; x <= xhi; ++x) {
            const float xg = ((float)x + 0.5f) * iz - o.x;
            const float dx = xg - G.x;
            if (dx * dx + dy2 <= G.z) {
                // the grain's cell must lie in the point's window
                if (floor_div(xg - rm, delta, rd, fast) <= gi && gi <= floor_div(xg + rm, delta, rd, fast) &&
                    floor_div(yg - rm, delta, rd, fast) <= gj && gj <= floor_div(yg + rm, delta, rd, fast)) {
                    atomicOr(&bits[((y - t.Y0) * t.TW + (x - t.X0)) * a.words + (k >> 5)], 1u << (k & 31));
                }
            }
        }
    }
}

// Cancellation between tiles (a workgroup reads the flag once per tile):
// true when the host has raised *flag; the caller's status then records it.
// The flag lives in coherent host memory, so the load goes to the system
// scope.  Must be called by every thread of the workgroup (it synchronises).
__device__ __forceinline__ bool tile_cancelled(const int *flag, int *status)
{
    __shared__ int s_cancel;
    if (!flag) return false;
    if (threadIdx.x == 0) {
        s_cancel = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (s_cancel) atomicOr(status, kStatusCancelled);
    }
    __syncthreads();
    return s_cancel != 0;
}

// Wave-wide inclusive sum.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// v of lane l (a per-lane index, 0 <= l < 64); ba = 4 l, the byte address
// ds_bpermute takes
__device__ __forceinline__ int from_lane_i(int v, int ba) { return __builtin_amdgcn_ds_bpermute(ba, v); }
// A wave-uniform float moved to an SGPR (no scalar float ALU on this target:
// a float computed from uniform values lands in a VGPR and holds it).
__device__ __forceinline__ float uniform_f(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ float from_lane_f(float v, int ba)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(ba, __float_as_int(v)));
}

}  // namespace fg

#endif
