// fg_batch.hip -- device helpers of the batched persistent renderer.
//
// Everything here is a plain kernel (no library calls, no memsets), so the
// whole batch pipeline can be captured in a HIP graph and replayed: the bin
// scan, the work of each (plane, tile) item, the heaviest-first item order,
// counter fills and the grain-capacity check.
#include <algorithm>

#include "fg_device.h"
#include "fg_pixelwise_hbm.h"
#include "fg_region.h"

namespace fg {

constexpr int kScanItems = 4;                       // elements per thread
constexpr int kScanTile = kBlock * kScanItems;      // elements per workgroup

size_t scan_partials(long long n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

// Workgroup sums of kScanTile-element tiles.
__global__ __launch_bounds__(kBlock) void scan_tile_sums(const uint32_t *in, long long n, uint32_t *partials)
{
    __shared__ uint32_t s_wave[kBlock / 64];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    uint32_t v = 0;
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) v += in[base + i];
    uint32_t total;
    (void)block_exclusive_scan(v, s_wave, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Exclusive scan of the tile sums in place (one workgroup, any count).
__global__ __launch_bounds__(kBlock) void scan_tile_prefix(uint32_t *partials, long long m)
{
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (long long b = 0; b < m; b += kBlock) {
        const long long i = b + threadIdx.x;
        const uint32_t v = i < m ? partials[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, s_wave, total);
        const uint32_t carry = s_carry;
        if (i < m) partials[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[m] = s_carry;
}

// out[i] = tile prefix + exclusive prefix within the tile; out[n] = total.
__global__ __launch_bounds__(kBlock) void scan_tile_apply(const uint32_t *in, uint32_t *out, long long n,
                                                          const uint32_t *partials, long long m)
{
    __shared__ uint32_t s_wave[kBlock / 64];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    uint32_t v[kScanItems], sum = 0;
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = base + i < n ? in[base + i] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t run = partials[blockIdx.x] + block_exclusive_scan(sum, s_wave, total);
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = partials[m];
}

hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, long long n, uint32_t *partials, hipStream_t s)
{
    const long long m = (n + kScanTile - 1) / kScanTile;
    if (m > 0) {
        hipLaunchKernelGGL(scan_tile_sums, dim3((unsigned)m), dim3(kBlock), 0, s, in, n, partials);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(scan_tile_prefix, dim3(1), dim3(kBlock), 0, s, partials, m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scan_tile_apply, dim3((unsigned)std::max<long long>(m, 1)), dim3(kBlock), 0, s, in, out, n,
                       partials, m);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void fill_u32_kernel(uint32_t *p, long long n, uint32_t v)
{
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        p[i] = v;
}

hipError_t launch_fill_u32(uint32_t *p, long long n, uint32_t v, hipStream_t s)
{
    const long long blocks = std::max<long long>(1, std::min<long long>((n + kBlock - 1) / kBlock, 4096));
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, n, v);
    return hipGetLastError();
}

// The plane of batch item `item` (planes sorted by item_base).
__device__ __forceinline__ int plane_of(const BatchPlane *planes, int n_planes, long long item)
{
    int lo = 0, hi = n_planes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (planes[mid].item_base <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t item_weight(const BatchPlane &p, int tile)
{
    const PixelArgs &a = p.a;
    const int tyi = tile / a.tiles_x, txi = tile - tyi * a.tiles_x;
    const int X0 = txi * a.tile_w, Y0 = tyi * a.tile_h;
    const TileRegion r = tile_region(a, p.geo, X0, Y0, min(X0 + a.tile_w, a.out_w), min(Y0 + a.tile_h, a.out_h));
    uint32_t sum = 0;
    if (r.bx0 <= r.bx1)
        for (int by = r.by0; by <= r.by1; ++by)
            sum += p.bin_start[by * p.geo.nbx + r.bx1 + 1] - p.bin_start[by * p.geo.nbx + r.bx0];
    return sum;
}

__global__ __launch_bounds__(kBlock) void batch_weights_kernel(const BatchPlane *planes, int n_planes, uint32_t *w,
                                                               int n_items)
{
    const int item = blockIdx.x * kBlock + threadIdx.x;
    if (item >= n_items) return;
    const int pl = plane_of(planes, n_planes, item);
    w[item] = item_weight(planes[pl], (int)(item - planes[pl].item_base));
}

hipError_t launch_batch_weights(const BatchPlane *planes, int n_planes, uint32_t *w, int n_items, hipStream_t s)
{
    if (n_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_weights_kernel, dim3((unsigned)((n_items + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       planes, n_planes, w, n_items);
    return hipGetLastError();
}

// Counting sort of the items by bucket floor(log2(w + 1)), heaviest bucket
// first (one workgroup; the order inside a bucket is whatever the atomics
// give, which changes no result, only the schedule).
__device__ __forceinline__ int weight_bucket(uint32_t w)
{
    return 32 - __clz(w == 0xFFFFFFFFu ? w : w + 1u);  // 1..32
}

__global__ __launch_bounds__(1024) void batch_order_kernel(const BatchPlane *planes, int n_planes, const uint32_t *w,
                                                           int n_items, int2 *items)
{
    __shared__ uint32_t s_count[64], s_pos[64];
    if (threadIdx.x < 64) s_count[threadIdx.x] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n_items; i += 1024) atomicAdd(&s_count[weight_bucket(w[i])], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = 63; b >= 0; --b) {
            s_pos[b] = run;
            run += s_count[b];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_items; i += 1024) {
        const uint32_t pos = atomicAdd(&s_pos[weight_bucket(w[i])], 1u);
        const int pl = plane_of(planes, n_planes, i);
        items[pos] = make_int2(pl, (int)(i - planes[pl].item_base));
    }
}

hipError_t launch_batch_order(const BatchPlane *planes, int n_planes, const uint32_t *w, int n_items, int2 *items,
                              uint32_t *scratch, hipStream_t s)
{
    (void)scratch;
    if (n_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_order_kernel, dim3(1), dim3(1024), 0, s, planes, n_planes, w, n_items, items);
    return hipGetLastError();
}

// *out += the sum of n u32 values in 64 bits (one atomic per workgroup).
__global__ __launch_bounds__(kBlock) void sum_u64_kernel(const uint32_t *in, long long n, unsigned long long *out)
{
    __shared__ unsigned long long s_part[kBlock / 64];
    unsigned long long acc = 0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        acc += in[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += s_part[w];
        if (t) atomicAdd(out, t);
    }
}

hipError_t launch_sum_u64(const uint32_t *in, long long n, unsigned long long *out, hipStream_t s)
{
    const long long blocks = std::max<long long>(1, std::min<long long>((n + kBlock - 1) / kBlock, 512));
    hipLaunchKernelGGL(sum_u64_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, in, n, out);
    return hipGetLastError();
}

__global__ void capacity_check_kernel(const unsigned long long *total, uint32_t cap, int *status)
{
    if (*total > cap) atomicOr(status, kStatusCapacity);
}

hipError_t launch_capacity_check(const unsigned long long *total, uint32_t cap, int *status, hipStream_t s)
{
    hipLaunchKernelGGL(capacity_check_kernel, dim3(1), dim3(1), 0, s, total, cap, status);
    return hipGetLastError();
}

}  // namespace fg
