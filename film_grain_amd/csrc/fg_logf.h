// fg_logf.h -- the f32 natural logarithm of the host C library (glibc's logf,
// which the reference's `f32::ln` calls on Linux: src/model.rs:252-265 takes
// ln(max(1 - u, 1e-6)) per pixel), reproduced bit for bit on the device for
// inputs in [1e-6, 1], the domain of the lambda pre-pass.
//
// glibc's logf is not correctly rounded (over the 1.67e8 floats of that
// domain it differs from the correctly rounded logarithm at about 1.8e5), so
// no accurate logarithm alone can reproduce it.  Instead:
//   fg_log_double(x)  ln x in double precision from IEEE +, -, *, / only
//                     (so host and device compute the same bits), via
//                     x = 2^k m, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s),
//                     s = (m - 1) / (m + 1), atanh by its Taylor series to
//                     s^23 (|s| < 0.172: truncation below 1e-20 relative),
//                     k ln 2 with ln 2 split hi + lo (fdlibm's split: hi has
//                     trailing zeros, so k * hi is exact);
//   table             every input of the domain where (float)fg_log_double(x)
//                     differs from the C library's logf(x), found by running
//                     both over the whole domain when the library is built
//                     (gen_logf_table.hip); each differs by one unit in the
//                     last place, stored as the input's bits with the
//                     direction in bit 31;
//   fg_logf_exact     the float of fg_log_double, corrected by a binary search
//                     of the table.
// The generator re-checks every input of the domain against logf after
// writing the table and fails the build on any difference.
#ifndef FG_LOGF_H
#define FG_LOGF_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

constexpr float kLogfLo = 1e-6f;  // domain [kLogfLo, 1]

__host__ __device__ inline uint32_t f32_bits(float x)
{
    union {
        float f;
        uint32_t u;
    } v;
    v.f = x;
    return v.u;
}

__host__ __device__ inline float bits_f32(uint32_t u)
{
    union {
        float f;
        uint32_t u;
    } v;
    v.u = u;
    return v.f;
}

// ln x for a positive normal float x (double precision, IEEE basic operations).
__host__ __device__ inline double fg_log_double(float x)
{
    const uint32_t b = f32_bits(x);
    int k = (int)(b >> 23) - 127;
    double m = (double)bits_f32((b & 0x007FFFFFu) | 0x3F800000u);  // [1, 2)
    if (m > 1.4142135623730951) {
        m = m * 0.5;  // exact
        k += 1;
    }
    // m - 1 is exact (Sterbenz) and m + 1 needs at most 26 bits
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    for (int n = 21; n >= 3; n -= 2) p = p * s2 + 1.0 / (double)n;
    p = p * s2 + 1.0;
    const double lnm = 2.0 * s * p;
    const double ln2_hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000
    const double ln2_lo = 1.90821492927058770002e-10;  // 0x3DEA39EF35793C76
    return (double)k * ln2_hi + ((double)k * ln2_lo + lnm);
}

// The table entry for an input: its bits, with bit 31 set when the library
// result is one unit above (in bit pattern) the float of fg_log_double.
__host__ __device__ inline float fg_logf_from_table(float x, const uint32_t *table, int n)
{
    const float base = (float)fg_log_double(x);
    const uint32_t key = f32_bits(x);
    int lo = 0, hi = n;  // first entry whose input is >= key
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((table[mid] & 0x7FFFFFFFu) < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < n && (table[lo] & 0x7FFFFFFFu) == key)
        return bits_f32(f32_bits(base) + ((table[lo] >> 31) ? 1u : 0xFFFFFFFFu));
    return base;
}

}  // namespace fg

#endif
