// fg_probe.hip -- the integrators' f64 math, element-wise, so that tests can
// compare the device's results with the C library's (fg_math_device).
#include <algorithm>

#include "fg_frame.h"
#include "fg_rng.h"

namespace fg {

constexpr int kProbeThreads = 256;

// The device's f64 math as the integrators evaluate it (fg_rng.h,
// fg_device.h), for comparison with the C library the reference calls.
__global__ __launch_bounds__(kProbeThreads) void math_probe_kernel(int fn, const double *in, long long n,
                                                                   double *out)
{
    const long long stride = (long long)gridDim.x * kProbeThreads;
    for (long long i = (long long)blockIdx.x * kProbeThreads + threadIdx.x; i < n; i += stride) {
        const double x = in[i];
        double r;
        switch (fn) {
        case kMathExpNeg: r = exp_neg_glibc((float)x); break;         // Knuth threshold (x: an f32 value)
        case kMathExp: r = fg_exp_dd(x); break;                       // rejection check, ziggurat
        case kMathLog: r = log(x); break;                             // rejection branch
        case kMathTanPi: r = tan(3.14159265358979323846 * x); break;  // Cauchy draw
        case kMathLogGamma: r = log_gamma(x); break;                  // rand_distr log_gamma
        case kMathSqrt: r = sqrt(x); break;
        case kMathExpF32: r = (double)exp_to_f32(x); break;           // LogNormal radius as f32
        default: r = 0.0; break;
        }
        out[i] = r;
    }
}

hipError_t launch_math_probe(int fn, const double *in, long long n, double *out, hipStream_t s)
{
    const long long blocks = std::min<long long>(std::max<long long>((n + kProbeThreads - 1) / kProbeThreads, 1),
                                                 65536);
    hipLaunchKernelGGL(math_probe_kernel, dim3((unsigned)blocks), dim3(kProbeThreads), 0, s, fn, in, n, out);
    return hipGetLastError();
}

}  // namespace fg
