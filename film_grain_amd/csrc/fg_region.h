// fg_region.h -- cell region of a pixel-wise output tile (device code shared
// by the splat kernels and the batch scheduler).
#ifndef FG_REGION_H
#define FG_REGION_H

#include "fg_device.h"
#include "fg_pixelwise_hbm.h"

namespace fg {

// Cell region of output rect [X0, X1) x [Y0, Y1) (the union of its points'
// windows, exact: every step from a pixel index to a window bound is
// monotone), relative to the extended grid, and the bins covering it.
struct TileRegion {
    int ci0, ci1, cj0, cj1;
    int bx0, bx1, by0, by1;
};
__device__ __forceinline__ TileRegion tile_region(const PixelArgs &a, const PixGeom &geo, int X0, int Y0, int X1,
                                                  int Y1)
{
    const float rm = a.rm, delta = a.delta, rd = a.rdelta, iz = a.inv_zoom;
    const int fast = a.div_fast;
    const float4 orange = *a.off_range;
    const float px_lo = ((float)X0 + 0.5f) * iz, px_hi = ((float)(X1 - 1) + 0.5f) * iz;
    const float py_lo = ((float)Y0 + 0.5f) * iz, py_hi = ((float)(Y1 - 1) + 0.5f) * iz;
    const int I0 = floor_div((px_lo - orange.y) - rm, delta, rd, fast);
    const int I1 = floor_div((px_hi - orange.x) + rm, delta, rd, fast);
    const int J0 = floor_div((py_lo - orange.w) - rm, delta, rd, fast);
    const int J1 = floor_div((py_hi - orange.z) + rm, delta, rd, fast);
    TileRegion r;
    r.bx0 = max(I0 - geo.Ix0, 0) / kBinSide;
    r.bx1 = min(I1 - geo.Ix0, geo.W - 1) / kBinSide;
    r.by0 = max(J0 - geo.Jy0, 0) / kBinSide;
    r.by1 = min(J1 - geo.Jy0, geo.H - 1) / kBinSide;
    r.ci0 = I0 - geo.Ix0;
    r.ci1 = I1 - geo.Ix0;
    r.cj0 = J0 - geo.Jy0;
    r.cj1 = J1 - geo.Jy0;
    return r;
}

}  // namespace fg

#endif
