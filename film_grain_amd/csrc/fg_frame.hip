// fg_frame.hip -- the per-frame work around the integrators, on the device:
// colour split (src/color.rs:158-213), plane normalisation and the lambda
// plane (src/model.rs:228-265), colour merge with the nearest resize of the
// chroma planes and the 8-bit quantisation (src/color.rs:66-114, 237-239;
// src/model.rs:77-98).  Every value is bit-identical to the host restatement
// (film_grain_amd/csrc/host): the same f32 operations in the same order,
// no contraction, and the C library's logf through fg_logf.h.
#include "fg_frame.h"
#include "fg_logf.h"
#include "fg_logf_table.h"

namespace fg {

namespace {

constexpr float kEps = 1e-6f;       // src/model.rs: EPSILON
constexpr float kMaxLambda = 1e6f;  // src/model.rs: MAX_LAMBDA
constexpr float kYR = 0.2126f, kYG = 0.7152f, kYB = 0.0722f;  // src/color.rs:9-11
constexpr float kCbDenom = 1.8556f, kCrDenom = 1.5748f;       // src/color.rs:12-13
constexpr int kThreads = 256;

__device__ __forceinline__ float clamp01(float v)  // f32::clamp(0, 1) keeps NaN
{
    if (v < 0.0f) return 0.0f;
    if (v > 1.0f) return 1.0f;
    return v;
}

__device__ __forceinline__ float max_nan(float a, float b)  // f32::max: a NaN operand loses
{
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}

__device__ __forceinline__ uint8_t to_u8(float v)  // `(clamp01(v) * 255 + 0.5).floor() as u8`
{
    const float f = floorf(clamp01(v) * 255.0f + 0.5f);
    if (!(f > 0.0f)) return 0;
    if (f >= 255.0f) return 255;
    return (uint8_t)f;
}

__device__ __forceinline__ unsigned grid_threads() { return gridDim.x * kThreads; }

}  // namespace

__device__ float fg_logf_exact(float x) { return fg_logf_from_table(x, kLogfTable, kLogfTableN); }

__global__ __launch_bounds__(kThreads) void log_kernel(const float *in, long long n, float *out)
{
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += grid_threads())
        out[i] = fg_logf_exact(in[i]);
}

// max(0, every non-NaN value) as float bits: the fold of normalize_plane
// starts at 0.0 with f32::max, so only positive values can raise it, and the
// bits of non-negative floats order as the floats do.
__global__ __launch_bounds__(kThreads) void plane_max_kernel(const float *in, long long n, unsigned *max_bits)
{
    unsigned m = 0;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += grid_threads()) {
        const float v = in[i];
        if (v > 0.0f) m = max(m, f32_bits(v));
    }
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(max_bits, m);
}

__global__ __launch_bounds__(kThreads) void lambda_kernel(const float *in, long long n, const unsigned *max_bits,
                                                          float inv_e_pi_r2, float *norm, float *lambda)
{
    const float mx = bits_f32(*max_bits);
    const bool already = mx <= 1.0f + kEps;
    const float denom = max_nan(mx + kEps, kEps);
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += grid_threads()) {
        float v = already ? in[i] : in[i] / denom;
        if (v < 0.0f) v = 0.0f;  // f32::clamp keeps NaN
        if (v > 1.0f - kEps) v = 1.0f - kEps;
        if (norm) norm[i] = v;
        // lambda_plane clamps again (a no-op here) and guards the logarithm
        const float safe = max_nan(1.0f - v, kEps);
        const float activity = -inv_e_pi_r2 * fg_logf_exact(safe);
        lambda[i] = activity < kMaxLambda ? activity : kMaxLambda;  // f32::min ignores NaN
    }
}

__global__ __launch_bounds__(kThreads) void split_kernel(int mode, const float *rgb, long long n, float *p0,
                                                         float *p1, float *p2)
{
    for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < n; k += grid_threads()) {
        const float r = clamp01(rgb[3 * k]), g = clamp01(rgb[3 * k + 1]), b = clamp01(rgb[3 * k + 2]);
        if (mode == kColorRgb) {
            p0[k] = r;
            p1[k] = g;
            p2[k] = b;
        } else {
            const float luma = kYR * r + kYG * g + kYB * b;
            p0[k] = clamp01(luma);
            p1[k] = (b - luma) / kCbDenom;
            p2[k] = (r - luma) / kCrDenom;
        }
    }
}

// Plane::resize_nearest's source index along one axis (src/model.rs:77-98).
__device__ __forceinline__ int nearest_src(int x, float scale, int src_n)
{
    const float f = fminf(fmaxf(((float)x + 0.5f) * scale - 0.5f, 0.0f), (float)(src_n - 1));
    return (int)roundf(f);
}

__global__ __launch_bounds__(kThreads) void merge_kernel(int mode, const float *p0, const float *p1,
                                                         const float *p2, int w, int h, int cw, int ch,
                                                         uint8_t *rgb8)
{
    const long long n = (long long)w * h;
    const float sx = (float)cw / (float)w, sy = (float)ch / (float)h;
    for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < n; k += grid_threads()) {
        float r, g, b;
        if (mode == kColorRgb) {
            r = p0[k];
            g = p1[k];
            b = p2[k];
        } else {
            const int y = (int)(k / w), x = (int)(k - (long long)y * w);
            // chroma planes keep the input size and are resized to the output
            const long long c = (cw == w && ch == h)
                                    ? k
                                    : (long long)nearest_src(y, sy, ch) * cw + nearest_src(x, sx, cw);
            const float yv = p0[k], cbv = p1[c], crv = p2[c];
            r = clamp01(yv + kCrDenom * crv);
            b = clamp01(yv + kCbDenom * cbv);
            g = clamp01((yv - kYR * r - kYB * b) / kYG);
        }
        rgb8[3 * k] = to_u8(r);
        rgb8[3 * k + 1] = to_u8(g);
        rgb8[3 * k + 2] = to_u8(b);
    }
}

__global__ void zero_u32_kernel(unsigned *p) { *p = 0u; }

static unsigned blocks_for(long long n)
{
    const long long b = (n + kThreads - 1) / kThreads;
    return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

hipError_t launch_log_f32(const float *in, long long n, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(log_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, s, in, n, out);
    return hipGetLastError();
}

hipError_t launch_lambda(const float *in, long long n, float inv_e_pi_r2, unsigned *max_bits, float *norm,
                         float *lambda, hipStream_t s)
{
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, max_bits);
    hipLaunchKernelGGL(plane_max_kernel, dim3(std::min(blocks_for(n), 2048u)), dim3(kThreads), 0, s, in, n,
                       max_bits);
    hipLaunchKernelGGL(lambda_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, s, in, n, max_bits, inv_e_pi_r2, norm,
                       lambda);
    return hipGetLastError();
}

hipError_t launch_split(int mode, const float *rgb, long long n, float *p0, float *p1, float *p2, hipStream_t s)
{
    hipLaunchKernelGGL(split_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, s, mode, rgb, n, p0, p1, p2);
    return hipGetLastError();
}

hipError_t launch_merge(int mode, const float *p0, const float *p1, const float *p2, int w, int h, int cw, int ch,
                        uint8_t *rgb8, hipStream_t s)
{
    hipLaunchKernelGGL(merge_kernel, dim3(blocks_for((long long)w * h)), dim3(kThreads), 0, s, mode, p0, p1, p2, w,
                       h, cw, ch, rgb8);
    return hipGetLastError();
}

}  // namespace fg
