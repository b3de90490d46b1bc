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
//   bin_grains  each bin's grains written contiguously at bin_start[bin]; the
//               non-empty cells are compacted first so the divergent grain
//               draws run on lanes that all have work
//   splat       per output tile: the grains of the bins covering the tile's
//               cell region are streamed from HBM into LDS batches and tested
//               per sample through candidate boxes and the exact reference test
//               (as the tile splat kernel does), into an LDS bitset.
// A grain record is (cx, cy, r, cell) with cell = (i - Ix0) | (j - Jy0) << 16.
#include <hipcub/hipcub.hpp>

#include "fg_device.h"
#include "fg_pixelwise_hbm.h"

namespace fg {

__global__ __launch_bounds__(kBlock) void threshold_kernel(const float *lambda, long long n, float delta,
                                                           double *T)
{
    const long long p = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (p < n) T[p] = cell_threshold(lambda[p], delta);
}

// Cell (i, j) of bin `bin`, thread t: returns false outside the extended grid.
__device__ __forceinline__ bool bin_cell(const PixGeom &g, int bin, int t, int &i, int &j, uint32_t &packed)
{
    const int by = bin / g.nbx, bx = bin - by * g.nbx;
    const int ci = bx * kBinSide + (t & (kBinSide - 1)), cj = by * kBinSide + (t / kBinSide);
    if (ci >= g.W || cj >= g.H) return false;
    i = g.Ix0 + ci;
    j = g.Jy0 + cj;
    packed = (uint32_t)ci | ((uint32_t)cj << 16);
    return true;
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

__global__ __launch_bounds__(kBlock) void bin_count_kernel(PixelArgs a, PixGeom g, const double *T,
                                                           uint32_t *bin_count)
{
    __shared__ uint32_t s_wave[kBlock / 64];
    const int bin = blockIdx.x;
    int i, j;
    uint32_t packed, q = 0;
    if (bin_cell(g, bin, threadIdx.x, i, j, packed)) {
        float sx, sy;
        const double e = cell_threshold_at(a, T, i, j, sx, sy);
        Xoshiro rng;
        q = cell_count(mix3(a.seed, kCellStream, i, j), e, rng);
    }
    uint32_t total;
    (void)block_exclusive_scan(q, s_wave, total);
    if (threadIdx.x == 0) bin_count[bin] = total;
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void bin_grains_kernel(PixelArgs a, PixGeom g, const double *T,
                                                            const uint32_t *bin_start, float4 *grains)
{
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t s_excl[kBlock];
    __shared__ uint16_t s_list[kBlock];
    const int bin = blockIdx.x, tid = threadIdx.x;
    int i = 0, j = 0;
    uint32_t packed = 0, q = 0;
    const bool valid = bin_cell(g, bin, tid, i, j, packed);
    if (valid) {
        float sx, sy;
        const double e = cell_threshold_at(a, T, i, j, sx, sy);
        Xoshiro rng;
        q = cell_count(mix3(a.seed, kCellStream, i, j), e, rng);
    }
    uint32_t total, nonempty;
    s_excl[tid] = block_exclusive_scan(q, s_wave, total);
    const uint32_t slot = block_exclusive_scan(q > 0 ? 1u : 0u, s_wave, nonempty);
    if (q > 0) s_list[slot] = (uint16_t)tid;
    __syncthreads();
    const uint32_t base = bin_start[bin];
    for (uint32_t e = tid; e < nonempty; e += kBlock) {
        const int t = s_list[e];
        int ci, cj;
        uint32_t pk;
        bin_cell(g, bin, t, ci, cj, pk);
        float sx, sy;
        const double th = cell_threshold_at(a, T, ci, cj, sx, sy);
        Xoshiro rng;
        const uint32_t n = cell_count(mix3(a.seed, kCellStream, ci, cj), th, rng);
        float4 *dst = grains + base + s_excl[t];
        const float pkf = __uint_as_float(pk);
        for (uint32_t k = 0; k < n; ++k) {
            const float cx = sx + uniform_f32(rng, a.uscale);
            const float cy = sy + uniform_f32(rng, a.uscale);
            float rad = a.r_const;
            if (KIND == 1) {
                rad = lognormal_f32(rng, a.log_mu, a.log_sigma, d_zig_x, d_zig_f);
                if (rad > a.rm) rad = a.rm;
            }
            dst[k] = make_float4(cx, cy, rad, pkf);
        }
    }
}

// ---------------------------------------------------------------------------
// splat from HBM grains
// ---------------------------------------------------------------------------
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void splat_hbm_kernel(PixelArgs a, PixGeom geo, const uint32_t *bin_start,
                                                          const float4 *grains)
{
    constexpr int NW = BLOCK / 64;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t s_queue[NW][kQueueCap];

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
    t.I0 = geo.Ix0;  // grain cells are packed relative to the extended grid
    t.J0 = geo.Jy0;
    const float rm = a.rm, delta = a.delta, rd = a.rdelta, iz = a.inv_zoom, z = a.zoom;
    const int fast = a.div_fast;
    const int N = a.n_samples, words = a.words;
    const float4 orange = *a.off_range;

    // cell region of the tile (exact), then the bins covering it
    const float px_lo = (t.fx0 + 0.5f) * iz, px_hi = (t.fx1 + 0.5f) * iz;
    const float py_lo = (t.fy0 + 0.5f) * iz, py_hi = (t.fy1 + 0.5f) * iz;
    const int I0 = floor_div((px_lo - orange.y) - rm, delta, rd, fast);
    const int I1 = floor_div((px_hi - orange.x) + rm, delta, rd, fast);
    const int J0 = floor_div((py_lo - orange.w) - rm, delta, rd, fast);
    const int J1 = floor_div((py_hi - orange.z) + rm, delta, rd, fast);
    const int bx0 = max(I0 - geo.Ix0, 0) / kBinSide, bx1 = min(I1 - geo.Ix0, geo.W - 1) / kBinSide;
    const int by0 = max(J0 - geo.Jy0, 0) / kBinSide, by1 = min(J1 - geo.Jy0, geo.H - 1) / kBinSide;
    // region in packed-cell coordinates, for culling grains of partly covered bins
    const int rci0 = I0 - geo.Ix0, rci1 = I1 - geo.Ix0, rcj0 = J0 - geo.Jy0, rcj1 = J1 - geo.Jy0;

    // ---- LDS layout ----
    size_t off = 0;
    uint32_t *bits = reinterpret_cast<uint32_t *>(lds);
    const int tile_words = a.tile_w * a.tile_h * words;
    off += ((size_t)tile_words * 4 + 15) & ~(size_t)15;
    const float2 *offs = a.offs;
    if (a.offs_in_lds) {
        float2 *ol = reinterpret_cast<float2 *>(lds + off);
        for (int k = tid; k < N; k += BLOCK) ol[k] = a.offs[k];
        offs = ol;
        off += ((size_t)N * 8 + 15) & ~(size_t)15;
    }
    const int gcap = (int)min((long long)((a.lds_bytes - (long long)off) / 32), 65535LL);
    if (gcap < BLOCK) {
        if (tid == 0) atomicOr(a.status, kStatusRegion);
        return;
    }
    float4 *gA = reinterpret_cast<float4 *>(lds + off);
    float4 *gB = gA + gcap;
    for (int w = tid; w < tile_words; w += BLOCK) bits[w] = 0u;

    // sample mapping: lanes over samples k, groups over grains
    const int KL = N >= BLOCK ? BLOCK : N;
    const int GS = BLOCK / KL;
    const int kl = tid % KL, gsub = tid / KL;
    const bool tester = tid < KL * GS;
    const int KM = (N + KL - 1) / KL;
    float oxz0 = 0.0f, oyz0 = 0.0f;
    // the wave's extent of scaled offsets, for whole-wave grain culling
    float ox_lo = INFINITY, ox_hi = -INFINITY, oy_lo = INFINITY, oy_hi = -INFINITY;
    if (tester) {
        for (int m = 0; m < KM; ++m) {
            const int k = kl + m * KL;
            if (k >= N) break;
            const float2 o = a.offs[k];
            const float ox = o.x * z, oy = o.y * z;
            if (m == 0) {
                oxz0 = ox;
                oyz0 = oy;
            }
            ox_lo = fminf(ox_lo, ox);
            ox_hi = fmaxf(ox_hi, ox);
            oy_lo = fminf(oy_lo, oy);
            oy_hi = fmaxf(oy_hi, oy);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        ox_lo = fminf(ox_lo, __shfl_xor(ox_lo, o, 64));
        ox_hi = fmaxf(ox_hi, __shfl_xor(ox_hi, o, 64));
        oy_lo = fminf(oy_lo, __shfl_xor(oy_lo, o, 64));
        oy_hi = fmaxf(oy_hi, __shfl_xor(oy_hi, o, 64));
    }
    const float slack = 4e-3f + 2e-6f * (float)max(a.out_w, a.out_h);
    uint32_t *queue = s_queue[wid];
    int qcount = 0;
    const bool one_k = KM == 1;  // N <= BLOCK: one sample per lane, offset held in registers
    const bool kvalid = tester && kl < N;
    const uint32_t klane = (uint32_t)kl << 16;

    // queue the (grain, sample) pairs whose candidate box meets the tile
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
            // whole-wave cull: no sample of this wave puts the grain's box in the tile
            const bool reach = gv && B.x + ox_hi + B.z >= t.fx0 && B.x + ox_lo - B.z <= t.fx1 &&
                               B.y + oy_hi + B.z >= t.fy0 && B.y + oy_lo - B.z <= t.fy1;
            if (!__any(reach)) continue;
            if (one_k) {
                const SplatBox b = splat_box(B, oxz0, oyz0);
                push(reach && kvalid && splat_box_hits(b, t), (uint32_t)g | klane);
            } else {
                for (int m = 0; m < KM; ++m) {
                    const int k = kl + m * KL;
                    const float2 o = offs[k < N ? k : 0];
                    const SplatBox b = splat_box(B, o.x * z, o.y * z);
                    push(reach && k < N && splat_box_hits(b, t), (uint32_t)g | ((uint32_t)k << 16));
                }
            }
        }
        if (qcount > 0) {
            splat_drain(a, t, queue, qcount, qcount, gA, gB, offs, bits);
            qcount = 0;
        }
        __syncthreads();
    };

    // ---- stream the region's grains through LDS batches ----
    const float zs = z;
    int fill = 0;  // uniform
    for (int by = by0; by <= by1; ++by) {
        uint32_t s = bin_start[by * geo.nbx + bx0];
        const uint32_t e = bin_start[by * geo.nbx + bx1 + 1];
        while (s < e) {
            const int n = (int)min((uint32_t)(gcap - fill), e - s);
            for (int q = tid; q < n; q += BLOCK) {
                const float4 G = grains[s + q];
                const uint32_t pk = __float_as_uint(G.w);
                const int ci = (int)(pk & 0xFFFFu), cj = (int)(pk >> 16);
                const bool inreg = ci >= rci0 && ci <= rci1 && cj >= rcj0 && cj <= rcj1;
                const float rad = G.z;
                // a grain outside the region can enter no window of the tile:
                // give it an empty box so it is culled
                const float r2 = (rad <= 0.0f || !inreg) ? -1.0f : rad * rad;
                const float rz = inreg ? fmaxf(rad, 0.0f) * zs + slack : -1.0e30f;
                gA[fill + q] = make_float4(G.x, G.y, r2, G.w);
                gB[fill + q] = make_float4(G.x * zs - 0.5f, G.y * zs - 0.5f, rz, 0.0f);
            }
            fill += n;
            s += (uint32_t)n;
            if (fill == gcap) {
                flush(fill);
                fill = 0;
            }
        }
    }
    if (fill > 0) flush(fill);
    __syncthreads();

    // ---- popcount epilogue: one store per output pixel ----
    const int tw = t.X1 - t.X0, th = t.Y1 - t.Y0;
    for (int p = tid; p < tw * th; p += BLOCK) {
        const int ox = p % tw, oy = p / tw;
        const uint32_t *pw = bits + ((size_t)oy * a.tile_w + ox) * words;
        uint32_t count = 0;
        for (int w = 0; w < words; ++w) count += __popc(pw[w]);
        a.out[(size_t)(t.Y0 + oy) * a.out_w + (t.X0 + ox)] = (float)count * a.inv_samples;
    }
}

hipError_t launch_threshold(const PixelArgs &a, double *T, hipStream_t s)
{
    const long long n = (long long)a.in_w * a.in_h;
    hipLaunchKernelGGL(threshold_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       a.lambda, n, a.delta, T);
    return hipGetLastError();
}

hipError_t launch_bin_count(const PixelArgs &a, const PixGeom &g, const double *T, uint32_t *bin_count,
                            hipStream_t s)
{
    hipLaunchKernelGGL(bin_count_kernel, dim3(g.nbx * g.nby), dim3(kBlock), 0, s, a, g, T, bin_count);
    return hipGetLastError();
}

hipError_t scan_bins(void *temp, size_t *temp_bytes, const uint32_t *in, uint32_t *out, int n, hipStream_t s)
{
    return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, in, out, n, s);
}

hipError_t launch_bin_grains(const PixelArgs &a, const PixGeom &g, int dist_kind, const double *T,
                             const uint32_t *bin_start, float4 *grains, hipStream_t s)
{
    if (dist_kind == 0)
        hipLaunchKernelGGL((bin_grains_kernel<0>), dim3(g.nbx * g.nby), dim3(kBlock), 0, s, a, g, T, bin_start,
                           grains);
    else
        hipLaunchKernelGGL((bin_grains_kernel<1>), dim3(g.nbx * g.nby), dim3(kBlock), 0, s, a, g, T, bin_start,
                           grains);
    return hipGetLastError();
}

hipError_t launch_splat_hbm(const PixelArgs &a, const PixGeom &g, const uint32_t *bin_start,
                            const float4 *grains, int block, hipStream_t s)
{
    const dim3 grid(a.n_tiles);
    const size_t lds = (size_t)a.lds_bytes;
    if (block == 512)
        hipLaunchKernelGGL((splat_hbm_kernel<512>), grid, dim3(512), lds, s, a, g, bin_start, grains);
    else
        hipLaunchKernelGGL((splat_hbm_kernel<256>), grid, dim3(256), lds, s, a, g, bin_start, grains);
    return hipGetLastError();
}

}  // namespace fg
