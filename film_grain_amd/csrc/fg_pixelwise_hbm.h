// fg_pixelwise_hbm.h -- launchers of the per-plane grain generation and the
// splat-from-HBM pixel-wise kernel (fg_pixelwise_hbm.hip).
#ifndef FG_PIXELWISE_HBM_H
#define FG_PIXELWISE_HBM_H

#include "fg_kernels.h"

namespace fg {

constexpr int kBinSide = 16;  // a bin is 16 x 16 cells (one workgroup, one cell per thread)

// The plane's extended cell grid: every cell some point's window can reach.
struct PixGeom {
    int Ix0, Jy0;  // first cell
    int W, H;      // cells per row / column (each <= 65535)
    int nbx, nby;  // bins per row / column
};

hipError_t launch_threshold(const PixelArgs &a, double *T, hipStream_t s);
hipError_t launch_bin_count(const PixelArgs &a, const PixGeom &g, const double *T, uint32_t *bin_count,
                            hipStream_t s);
// hipCUB exclusive sum; with temp == nullptr only sets *temp_bytes.
hipError_t scan_bins(void *temp, size_t *temp_bytes, const uint32_t *in, uint32_t *out, int n, hipStream_t s);
hipError_t launch_bin_grains(const PixelArgs &a, const PixGeom &g, int dist_kind, const double *T,
                             const uint32_t *bin_start, float4 *grains, hipStream_t s);
hipError_t launch_splat_hbm(const PixelArgs &a, const PixGeom &g, const uint32_t *bin_start,
                            const float4 *grains, int block, hipStream_t s);

}  // namespace fg

#endif
