// fg_frame.h -- launchers of the frame kernels (fg_frame.hip): colour split
// and merge, plane normalisation and lambda, and the exact f32 logarithm.
#ifndef FG_FRAME_H
#define FG_FRAME_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

constexpr int kColorLuma = 0;  // render Y; carry BT.709 Cb, Cr at the input size
constexpr int kColorRgb = 1;   // render R, G, B

// The integrators' f64 math on the device, element-wise (fn: the kMath*
// codes, equal to FG_MATH_* of the public header).
constexpr int kMathExpNeg = 0, kMathExp = 1, kMathLog = 2, kMathTanPi = 3, kMathLogGamma = 4, kMathSqrt = 5,
              kMathExpF32 = 6;
hipError_t launch_math_probe(int fn, const double *in, long long n, double *out, hipStream_t s);
// out[i] = the C library's logf(in[i]) for in[i] in [1e-6, 1].
hipError_t launch_log_f32(const float *in, long long n, float *out, hipStream_t s);
// normalize_plane then lambda_plane of n values; norm may be null; max_bits
// is one device word of scratch (cleared by the launch).
hipError_t launch_lambda(const float *in, long long n, float inv_e_pi_r2, unsigned *max_bits, float *norm,
                         float *lambda, hipStream_t s);
// Interleaved RGB f32 (n pixels) to three planes: R, G, B (clamped to [0, 1])
// or Y (clamped), Cb, Cr.
hipError_t launch_split(int mode, const float *rgb, long long n, float *p0, float *p1, float *p2, hipStream_t s);
// Three planes to interleaved 8-bit RGB of w x h: RGB planes all w x h; luma:
// Y is w x h, Cb and Cr are cw x ch and resized to w x h (nearest).
hipError_t launch_merge(int mode, const float *p0, const float *p1, const float *p2, int w, int h, int cw, int ch,
                        uint8_t *rgb8, hipStream_t s);

}  // namespace fg

#endif
