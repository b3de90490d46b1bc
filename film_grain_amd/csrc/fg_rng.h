// fg_rng.h -- random streams of the film-grain model, for host and gfx950 device.
//
// Restates, as __host__ __device__ code, the random stack the reference's CPU
// integrators draw from, so device results equal the CPU path on the same seed:
//   * cell / pixel hashing: src/rng.rs:26-52 (splitmix64, mix3, stream constants)
//   * rand 0.8.5 SmallRng = Xoshiro256++ seeded by four splitmix64 outputs;
//     next_u32 is the high half of next_u64
//   * rand 0.8.5 UniformFloat<f32> (23-bit mantissa, value*scale + low) and
//     Standard<f64> (53 high bits * 2^-53), Open01<f64>
//   * rand_distr 0.4.3 Poisson (Knuth below 12, Cauchy rejection above) with its
//     Lanczos log_gamma, StandardNormal ziggurat (256 layers), LogNormal
// All arithmetic keeps the reference's operation order and is compiled with
// -ffp-contract=off: Rust never fuses a*b+c.  The reference's exponentials are
// glibc's exp; on the device they are fg_exp_dd (fg_exp.h), which equals glibc
// but for about 2 inputs in 10^4, corrected by a table on the Knuth
// threshold's domain (fg_exp_glibc.h).  Host builds call glibc itself.
#ifndef FG_RNG_H
#define FG_RNG_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FG_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define FG_HD inline
#endif

#include "fg_zig_tables.h"
#if defined(__HIPCC__)
#include "fg_exp_glibc.h"
#endif

namespace fg {
// a <= exp(t) (or a < exp(t) with strict set) as the reference decides it
// with glibc's exp.  On the device the library's exp is within one unit of
// glibc's, so it decides unless a lies within a few units of it; there
// fg_exp_dd, glibc's but for its rare misroundings, decides.  (Using fg_exp_dd
// throughout costs its table loads in every wave that has one such lane.)
FG_HD bool cmp_exp(double a, double t, double scale, bool strict)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double e = scale * exp(t);
    if (fabs(a - e) <= 1.0e-15 * fabs(e)) e = scale * fg_exp_dd(t);
#else
    const double e = scale * exp(t);
#endif
    return strict ? a < e : a <= e;
}
}  // namespace fg

namespace fg {

constexpr uint64_t kOffsetStream = 0x9E3779B97F4A7C15ULL;  // src/rng.rs:5
constexpr uint64_t kCellStream = 0xA24B1C30BEBCCF59ULL;    // src/rng.rs:6
constexpr uint64_t kPixelStream = 0x6935FA5C55F65F1BULL;   // src/rng.rs:7
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

FG_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// src/rng.rs:46-52
FG_HD uint64_t splitmix64(uint64_t x)
{
    x += kGolden;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// src/rng.rs:36-38
FG_HD uint64_t mix(uint64_t seed, uint64_t stream) { return splitmix64(seed ^ stream); }

// src/rng.rs:40-44; a and b are i32 cell/pixel indices sign-extended to 64 bits.
FG_HD uint64_t mix3(uint64_t seed, uint64_t stream, int32_t a, int32_t b)
{
    uint64_t s = seed ^ stream;
    s = splitmix64(rotl64(s + (uint64_t)(int64_t)a, 17));
    return splitmix64(rotl64(s + (uint64_t)(int64_t)b, 41));
}

FG_HD float bits_f32(uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(b);
#else
    float f;
    memcpy(&f, &b, 4);
    return f;
#endif
}
FG_HD uint32_t f32_bits(float f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(f);
#else
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
#endif
}
FG_HD double bits_f64(uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __longlong_as_double((long long)b);
#else
    double d;
    memcpy(&d, &b, 8);
    return d;
#endif
}

// rand 0.8.5 Xoshiro256PlusPlus (SmallRng on 64-bit targets).
struct Xoshiro {
    uint64_t s0, s1, s2, s3;

    // SeedableRng::seed_from_u64: the state words are four consecutive
    // splitmix64 outputs starting from h; an all-zero state reseeds from 0.
    FG_HD void seed(uint64_t h)
    {
        s0 = splitmix64(h);
        s1 = splitmix64(h + kGolden);
        s2 = splitmix64(h + 2 * kGolden);
        s3 = splitmix64(h + 3 * kGolden);
        if ((s0 | s1 | s2 | s3) == 0) {
            uint64_t z = 0;
#if defined(__HIP_DEVICE_COMPILE__)
            // z stays opaque: folded to constants, the reseed's four words
            // were hoisted out of this branch (taken with probability
            // 2^-256) and spilled to scratch by every wave of the count
            // kernels
            asm volatile("" : "+v"(z));
#endif
            s0 = splitmix64(z);
            s1 = splitmix64(z + kGolden);
            s2 = splitmix64(z + 2 * kGolden);
            s3 = splitmix64(z + 3 * kGolden);
        }
    }

    FG_HD uint64_t next()
    {
        const uint64_t r = rotl64(s0 + s3, 23) + s0;
        const uint64_t t = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = rotl64(s3, 45);
        return r;
    }

    FG_HD uint32_t next_u32() { return (uint32_t)(next() >> 32); }
};

// UniformFloat<f32>::new(low, high): largest scale with low + scale*(1-2^-23) < high.
FG_HD float uniform_f32_scale(float low, float high)
{
    const float max_rand = bits_f32(0x3FFFFFFFu) - 1.0f;
    float scale = high - low;
    while (scale * max_rand + low >= high) scale = bits_f32(f32_bits(scale) - 1u);
    return scale;
}

// UniformFloat<f32>::sample with low = 0 (the only low the integrators use):
// (mantissa in [1,2) - 1) * scale + 0.
FG_HD float uniform_f32(Xoshiro &rng, float scale)
{
    const float v01 = bits_f32((rng.next_u32() >> 9) | 0x3F800000u) - 1.0f;
    return v01 * scale + 0.0f;
}

FG_HD double standard_f64(uint64_t v) { return (double)(v >> 11) * (1.0 / 9007199254740992.0); }
FG_HD double open01_f64(uint64_t v)
{
    return bits_f64(0x3FF0000000000000ULL | (v >> 12)) - (1.0 - 1.1102230246251565e-16);
}

// rand_distr 0.4.3 utils::log_gamma (Lanczos, g = 5, six terms).
FG_HD double log_gamma(double x)
{
    const double tmp = x + 5.5;
    const double lg = (x + 0.5) * log(tmp) - tmp;
    double a = 1.000000000190015;
    double d = x;
    d = d + 1.0; a = a + (76.18009172947146 / d);
    d = d + 1.0; a = a + (-86.50532032941677 / d);
    d = d + 1.0; a = a + (24.01409824083091 / d);
    d = d + 1.0; a = a + (-1.231739572450155 / d);
    d = d + 1.0; a = a + (0.1208650973866179e-2 / d);
    d = d + 1.0; a = a + (-0.5395239384953e-5 / d);
    return lg + log(2.5066282746310005 * a / x);
}

// Poisson<f64>::sample for lambda >= 12 (Cauchy comparison rejection), as f64.
FG_HD double poisson_rejection(Xoshiro &rng, double lambda)
{
    const double log_lambda = log(lambda);
    const double sqrt_2lambda = sqrt(2.0 * lambda);
    const double magic = lambda * log_lambda - log_gamma(1.0 + lambda);
    double result;
    for (;;) {
        double comp;
        for (;;) {
            comp = 0.0 + 1.0 * tan(3.14159265358979323846 * standard_f64(rng.next()));
            result = sqrt_2lambda * comp + lambda;
            if (result >= 0.0) break;
        }
        result = floor(result);
        // check = 0.9 (1 + comp^2) exp(...), and the draw is accepted when
        // U <= check
        if (cmp_exp(standard_f64(rng.next()), result * log_lambda - log_gamma(1.0 + result) - magic,
                    0.9 * (1.0 + comp * comp), false))
            break;
    }
    return result;
}

// Poisson<f64>::new(lambda).sample(rng) as u32.  Only exp(-lambda) is needed on
// the Knuth branch; the other precomputed terms of Poisson::new feed the
// rejection branch only.
FG_HD uint32_t poisson_u32(Xoshiro &rng, double lambda)
{
    if (lambda < 12.0) {
#if defined(__HIP_DEVICE_COMPILE__)
        const double e = exp_neg_glibc((float)lambda);  // lambda is an f32 value
#else
        const double e = exp(-lambda);
#endif
        double p = 1.0;
        uint32_t n = 0;
        while (p > e) {
            p = p * standard_f64(rng.next());
            ++n;
        }
        return n == 0 ? 0u : n - 1u;
    }
    const double r = poisson_rejection(rng, lambda);
    if (!(r > 0.0)) return 0u;
    if (r >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)r;
}

// StandardNormal via the 256-layer ziggurat of rand_distr 0.4.3 (symmetric
// variant: 8 low bits pick the layer, the high 52 bits give u in [-1, 1)).
template <class NextU64>
FG_HD double standard_normal(NextU64 &&next, const double *zx, const double *zf)
{
    for (;;) {
        const uint64_t bits = next();
        const int i = (int)(bits & 0xff);
        const double u = bits_f64((uint64_t)(1 + 1023) << 52 | (bits >> 12)) - 3.0;
        const double x = u * zx[i];
        if (fabs(x) < zx[i + 1]) return x;
        if (i == 0) {
            double xx = 1.0, yy = 0.0;
            while (-2.0 * yy < xx * xx) {
                const double x_ = open01_f64(next());
                const double y_ = open01_f64(next());
                xx = log(x_) / FG_ZIG_NORM_R;
                yy = log(y_);
            }
            return u < 0.0 ? xx - FG_ZIG_NORM_R : FG_ZIG_NORM_R - xx;
        }
        if (cmp_exp(zf[i + 1] + (zf[i] - zf[i + 1]) * standard_f64(next()), -x * x / 2.0, 1.0, true)) return x;
    }
}

// (float)exp(t) as the reference rounds glibc's exp(t) to f32.  On the device
// the library's exp is within one unit of glibc's, so the two round to the
// same f32 unless the double lies within a few units of an f32 rounding
// boundary (its 29 bits below the f32 mantissa near 0x10000000: about one
// draw in 10^8); there fg_exp_dd, glibc's but for its rare misroundings,
// decides.
FG_HD float exp_to_f32(double t)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double e = exp(t);
    const uint32_t low = (uint32_t)((uint64_t)__double_as_longlong(e) & 0x1FFFFFFFull);
    if (low - 0x0FFFFFF8u <= 16u) return (float)fg_exp_dd(t);
    return (float)e;
#else
    return (float)exp(t);
#endif
}

// LogNormal(mu, sigma).sample(rng) as f32 (src/model.rs:137-148).
FG_HD float lognormal_f32(Xoshiro &rng, double mu, double sigma, const double *zx, const double *zf)
{
    const double z = standard_normal([&rng]() { return rng.next(); }, zx, zf);
    return exp_to_f32(mu + sigma * z);
}

// Rust float->int `as` casts saturate and send NaN to 0.
FG_HD int32_t sat_i32(float v)
{
    if (!(v == v)) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)v;
}

}  // namespace fg

#endif
