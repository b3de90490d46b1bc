// fg_exp.h -- an exponential that host and device evaluate bit for bit alike,
// close enough to correctly rounded that it equals the C library's exp
// (glibc, which the reference's f64::exp calls) except on about 2 inputs in
// 10^4; fg_exp_neg_table.h lists those inputs for the Knuth threshold's
// domain, so that the device reproduces glibc's exp(-x) exactly there.
//
// fg_exp_dd(x): x = k ln2/64 + r with k = rint(x 64/ln2) and |r| <= ln2/128,
// exp(x) = 2^(k >> 6) 2^((k & 63)/64) exp(r).  2^(j/64) is a double-double
// (gen_exp_tables.py), r is carried as a double-double, exp(r) = 1 + r + q with
// q = r^2/2 + ... + r^7/5040 in double, and the product is formed with one
// exact product (fma) and summed with one final rounding: about 2^-67
// relative error before that rounding.  Only IEEE +, -, *, fma and rint are
// used, all correctly rounded on x86-64 and gfx950, so both give the same
// bits (compiled with -ffp-contract=off).
#ifndef FG_EXP_H
#define FG_EXP_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FG_EXP_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define FG_EXP_HD inline
#endif

#include "fg_exp_tables.h"

namespace fg {

#if defined(__HIPCC__)
static __constant__ double d_exp_t_hi[64] = {FG_EXP_T_HI_VALUES};
static __constant__ double d_exp_t_lo[64] = {FG_EXP_T_LO_VALUES};
#endif
static const double h_exp_t_hi[64] = {FG_EXP_T_HI_VALUES};
static const double h_exp_t_lo[64] = {FG_EXP_T_LO_VALUES};

FG_EXP_HD double exp_pow2i(int m)  // 2^m for -1022 <= m <= 1023
{
    const uint64_t b = (uint64_t)(1023 + m) << 52;
    double d;
    memcpy(&d, &b, 8);
    return d;
}

FG_EXP_HD double fg_exp_dd(double x)
{
    // NaN, infinities and results near the range's ends: the platform's exp
    if (!(x >= -708.0 && x <= 709.0)) return exp(x);
    const double kd = rint(x * FG_EXP_INV_L);
    const int k = (int)kd;
    // r = x - k L with L = ln2/64 = L_HI + L_LO: k L_HI is exact and so is
    // x - k L_HI (the two lie within a factor 2 of each other, or k = 0)
    const double rh = x - kd * FG_EXP_L_HI;
    const double p = kd * FG_EXP_L_LO;
    const double pe = fma(kd, FG_EXP_L_LO, -p);
    const double r = rh - p;  // two-sum of rh and -p
    const double bv = r - rh;
    const double rl = ((rh - (r - bv)) + (-p - bv)) - pe;
    const double q = r * r * (0.5 + r * (1.0 / 6.0 + r * (1.0 / 24.0 + r * (1.0 / 120.0 + r * (1.0 / 720.0 +
                                                                                          r * (1.0 / 5040.0)))))) +
                     r * rl;
    // s = 1 + r + (rl + q) as sh + sl
    const double sh = 1.0 + r;
    const double sb = sh - 1.0;
    const double sl = ((1.0 - (sh - sb)) + (r - sb)) + (rl + q);
    const int j = k & 63;
    const int m = (k - j) / 64;
#if defined(__HIP_DEVICE_COMPILE__)
    const double th = d_exp_t_hi[j], tl = d_exp_t_lo[j];
#else
    const double th = h_exp_t_hi[j], tl = h_exp_t_lo[j];
#endif
    const double ph = th * sh;
    const double pl = fma(th, sh, -ph);
    const double y = ph + ((pl + th * sl) + tl * sh);
    return y * exp_pow2i(m);
}

// exp(-x) as the C library computes it, for f32 x in (0, 12) (the Knuth
// threshold Poisson::new(x) takes): fg_exp_dd corrected by one unit in the
// last place at the inputs `table` lists (ascending f32 bits; bit 31 set: the
// library's result is one unit above).  Other x: fg_exp_dd alone.
FG_EXP_HD double fg_exp_neg_f32(float x, const uint32_t *table, int n)
{
    double e = fg_exp_dd(-(double)x);
    uint32_t b;
    memcpy(&b, &x, 4);
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((table[mid] & 0x7FFFFFFFu) < b)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < n && (table[lo] & 0x7FFFFFFFu) == b) {
        uint64_t eb;
        memcpy(&eb, &e, 8);
        eb = (table[lo] >> 31) ? eb + 1 : eb - 1;
        memcpy(&e, &eb, 8);
    }
    return e;
}

}  // namespace fg

#endif
