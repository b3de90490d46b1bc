// fg_kernels.h -- launch arguments and launchers of the gfx950 film-grain kernels.
// Internal to libfilmgrain_hip.so; the public boundary is include/film_grain_hip.h.
#ifndef FG_KERNELS_H
#define FG_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kMaxTilePixels = 1024;  // largest output tile one workgroup owns

// Device status word bits (OR-ed by kernels, read back by the host).
constexpr int kStatusNeighbourhood = 1;  // one output pixel's cell window exceeds LDS
constexpr int kStatusCancelled = 2;

// Pixel-wise integrator launch (src/pixelwise.rs:11-106).
struct PixelArgs {
    const float *lambda;        // in_h x in_w, row-major
    const float2 *offs;         // offsets_input, n_samples entries
    float *out;                 // out_h x out_w
    int *status;                // device status word
    unsigned long long *stats;  // optional counters: points, tests (T), hits
    const int *tile_list;       // optional tile indices (batching / resubmission)
    uint64_t seed;
    double log_mu, log_sigma;
    int in_w, in_h, out_w, out_h, n_samples;
    int tile_w, tile_h, tiles_x, n_tiles;
    int lds_bytes;
    float inv_zoom, rm, delta, uscale, r2_const, inv_samples;
    const float4 *off_range;    // device: (min x, max x, min y, max y) of offs
};

// Grain-wise integrator launch (src/grainwise.rs:12-124).
struct GrainArgs {
    const float *lambda;
    const float2 *offs;  // offsets in output pixels, n_samples entries
    float *out;
    int *status;
    unsigned long long *stats;  // optional counters: grains, (grain, sample) pairs, hits
    uint64_t seed;
    double log_mu, log_sigma;
    int in_w, in_h, out_w, out_h, n_samples, words;
    int tile_w, tile_h, tiles_x, n_tiles;
    int grains_per_thread;
    int lds_bytes;
    float zoom, rm, r_const, inv_samples, r_out_max;
    const float4 *off_range;    // device: (min x, max x, min y, max y) of offs
};

// Writes (min x, max x, min y, max y) of n (x, y) pairs to *range (one workgroup).
hipError_t launch_offsets_range(const float2 *offs, int n, float4 *range, hipStream_t s);
hipError_t launch_pixelwise(const PixelArgs &a, int dist_kind, bool with_stats, hipStream_t s);
hipError_t launch_grainwise(const GrainArgs &a, int dist_kind, bool with_stats, hipStream_t s);

}  // namespace fg

#endif
