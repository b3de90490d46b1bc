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
#include "fg_device.h"

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
    unsigned long long st_grains = 0, st_pairs = 0, st_hits = 0;
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
            if (lam > 0.0f) {
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
                        if (STATS && kl == 0) ++st_grains;
                        for (int k = kl; k < a.n_samples; k += KL) {
                            const float2 o = a.offs[k];
                            const float txp = G.x + o.x, typ = G.y + o.y;
                            if (STATS) ++st_pairs;
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
        atomicAdd(&a.stats[0], st_grains);
        atomicAdd(&a.stats[1], st_pairs);
        atomicAdd(&a.stats[2], st_hits);
    }
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
