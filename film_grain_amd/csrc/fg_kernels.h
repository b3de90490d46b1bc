// fg_kernels.h -- launch arguments and launchers of the gfx950 film-grain kernels.
// Internal to libfilmgrain_hip.so; the public boundary is include/film_grain_hip.h.
#ifndef FG_KERNELS_H
#define FG_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

constexpr int kBlock = 256;           // threads per workgroup (4 wave64)
constexpr int kMaxTilePixels = 1024;  // largest output tile of the gather kernel
constexpr int kQueueCap = 128;        // per-wave candidate queue of the splat kernel

// Device status word bits (OR-ed by kernels, read back by the host).
constexpr int kStatusNeighbourhood = 1;  // one output pixel's cell window exceeds LDS
constexpr int kStatusCancelled = 2;
constexpr int kStatusRegion = 4;         // a tile's cell region exceeds 65535 cells per side
constexpr int kStatusCapacity = 8;       // a batch's grains exceed the planned capacity
constexpr int kStatusLambda = 16;        // a lambda value is NaN, +inf or above kMaxLambda
constexpr int kStatusOverflow = 32;      // a plane's grain total reaches 2^32

// Largest lambda the integrators accept: lambda_plane caps every value at
// MAX_LAMBDA (src/model.rs:11, 262).  Values above it, +inf and NaN (on which
// Poisson::new fails in the reference) are skipped by the kernels, which never
// run a Poisson draw on them, and reported as kStatusLambda.
constexpr float kMaxLambda = 1.0e6f;

// Pixel-wise integrator launch (src/pixelwise.rs:11-106).
struct PixelArgs {
    const float *lambda;        // in_h x in_w, row-major
    const float2 *offs;         // offsets_input, n_samples entries
    float *out;                 // out_h x out_w
    int *status;                // device status word
    const int *cancel;          // optional: the context's host-visible cancel flag
    unsigned long long *stats;  // optional counters: points, tests (T), hits
    const int *tile_list;       // optional tile indices (batching / resubmission)
    const float4 *off_range;    // device: (min x, max x, min y, max y) of offs
    uint64_t seed;
    double log_mu, log_sigma;
    int in_w, in_h, out_w, out_h, n_samples;
    int tile_w, tile_h, tiles_x, n_tiles;
    int lds_bytes;
    int words;                  // ceil(n_samples / 32): bitset words per pixel (splat)
    int offs_in_lds;            // splat: stage the offsets table in LDS
    int div_fast;               // 1: floor((a)/delta) via the checked reciprocal sequence
    float inv_zoom, zoom, rm, delta, rdelta, uscale, r_const, r2_const, inv_samples;
    // batches: the grain total (device) and the grain array's capacity; when
    // the total exceeds it the grain and splat kernels do nothing (the
    // capacity check reports it).  grain_total == nullptr: unchecked.
    const unsigned long long *grain_total;
    uint32_t grain_cap;
    // adaptive sampling (HBM splat only): per output pixel its sample count
    // n in [1, n_samples] (the first n offsets, as a render with N = n uses),
    // per tile the index of the sample table covering its largest count, and
    // those tables (struct SampleTable, one per level).  budget == nullptr:
    // every pixel uses all n_samples, and level_tables holds the one table.
    const uint16_t *budget;
    const uint8_t *tile_level;
    const void *level_tables;
    // HBM splat, optional (cl_q == 0: off): candidate lists of the sorted
    // sample order per (radius class, cell of the grain's centre fractions),
    // for narrow grains that take the direct disk test; the lists of each
    // sample table are SampleTable::cl_cell / cl_list.  Class k holds grains
    // with rz < (k + 1) / cl_inv_dz; list (k, c) is cl_cell[k * cl_q^2 + c]
    // = (start, length) in cl_list.
    int cl_q, cl_classes;
    float cl_inv_dz;
};

// Grain-wise integrator launch (src/grainwise.rs:12-124).
struct GrainArgs {
    const float *lambda;
    const float2 *offs;  // offsets in output pixels, n_samples entries
    float *out;
    int *status;
    const int *cancel;   // optional: the context's host-visible cancel flag
    unsigned long long *stats;  // optional counters: grains, (grain, sample) pairs, disk tests, hits
    const float4 *off_range;    // device: (min x, max x, min y, max y) of offs
    uint64_t seed;
    double log_mu, log_sigma;
    int in_w, in_h, out_w, out_h, n_samples, words;
    int tile_w, tile_h, tiles_x, n_tiles;
    int grains_per_thread;
    int lds_bytes;
    float zoom, rm, r_const, inv_samples, r_out_max;
    // per input pixel, written by the Poisson pre-pass of the sorted kernel:
    // the grain count and the Xoshiro state after the count draws
    uint32_t *pix_q;
    ulonglong2 *pix_s01, *pix_s23;
    const int *tile_list;  // optional launch order of the tiles (sorted kernel)
    // optional candidate lists of the sorted kernel (constant radius): cell
    // qb * cell_q + qa of a cell_q x cell_q grid over the grain's fractional
    // target (A, B) lists sorted positions cell_list[cell_start[c] ..
    // cell_start[c + 1]); cell_q = 0: the per-grain run instead
    const uint32_t *cell_start;
    const uint16_t *cell_list;
    int cell_q;
};

// Writes (min x, max x, min y, max y) of n (x, y) pairs to *range (one workgroup).
hipError_t launch_offsets_range(const float2 *offs, int n, float4 *range, hipStream_t s);
// variant: 0 splat (default), 1 gather.  Stats always use the gather kernel.
hipError_t launch_pixelwise(const PixelArgs &a, int dist_kind, int variant, bool with_stats,
                            hipStream_t s);
hipError_t launch_grainwise(const GrainArgs &a, int dist_kind, bool with_stats, hipStream_t s);
// Instrumented twin: stats[0] += grains drawn with a positive output radius,
// stats[1] += those grains * N (the reference's (grain, sample) pairs).
hipError_t launch_grain_count(const GrainArgs &a, int dist_kind, hipStream_t s);

// Host check that q0 = a*rd, r = fma(-q0, d, a), q = fma(r, rd, q0) equals the
// IEEE quotient a/d for every a of one binade (hence, by scaling, for every a
// whose intermediates stay normal).  Cached per d by the caller.
bool reciprocal_division_exact(float d, float rd);

}  // namespace fg

#endif
