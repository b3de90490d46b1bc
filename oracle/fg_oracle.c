/*
 * fg_oracle.c -- CPU oracle for the film-grain hot path (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of the reference's CPU path.  Only tests/, the smoke
 * check and the benchmark's CPU-baseline leg use it, and only as a checker or
 * as the timed CPU baseline; the HIP product never calls it.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp).  Rust never
 * contracts a*b+c into an FMA and computes f32 arithmetic in single precision,
 * so -ffp-contract=off on x86-64 SSE reproduces its rounding exactly.
 *
 * Pinning (how each function is checked, see tests/test_oracle_kats.py):
 *   splitmix64/mix/mix3     closed form in src/rng.rs:36-52; published splitmix64
 *                           outputs for x=0.
 *   xoshiro256++            rand_xoshiro reference outputs for state [1,2,3,4].
 *   seed_from_u64           equals the splitmix64 sequence (closed form).
 *   UniformFloat<f32>       exhaustive check over all 2^23 mantissas (range, scale).
 *   Standard/Open01 f64,    driven with chosen next_u64 values: monotone in the
 *   UniformFloat sample     source, on the 2^-53 (2^-23 * scale) lattice, odd
 *                           multiples of 2^-53 for Open01, and independent of the
 *                           low bits each one discards (tests/test_oracle_bits.py).
 *   log_gamma               against libm lgamma (a wrong coefficient fails).
 *   Poisson                 mean/variance and chi-square against the pmf for both
 *                           branches (Knuth < 12, Cauchy rejection >= 12).
 *   ziggurat tables         recurrence invariants and the decimal literals of the
 *                           published table; StandardNormal moments and KS test.
 *   ChaCha                  RFC 7539 A.1 keystream vectors (20 rounds).
 *   PCG32 step and output   the published pcg32 demo vector (seed 42, stream
 *                           54); rand_core's increment constant is unpinned.
 *   inverse normal cdf      published quantiles; Phi(Phi^-1(p)) = p.
 *   pixel-wise / grain-wise brute force on tiny images (every grain enumerated
 *                           without the window), coverage invariant E[out] = u on
 *                           flat fields, grain-wise/pixel-wise agreement of means.
 *   cell / pixel dumps,     the same generator code as the integrators; digests
 *   digests, counts         and counts equal the dumps cell by cell
 *                           (tests/test_cellcheck.py, planted slips fail).
 *   colour split / merge,   BT.709's defining identities and closed forms on
 *   to_u8, resize_nearest   pure colours, split-then-merge exact on all 2^24
 *                           8-bit colours, to_u8 on every level, integer-zoom
 *                           replication and half-away rounding of the resize
 *                           (tests/test_oracle_colour.py).
 *   glibc_math              the C library itself (the reference's exp, log, tan).
 * Parity unpinned (no import of the Rust crates is possible here): the exact
 * decimal literals of rand_distr's ZIG_NORM tables beyond the entries checked,
 * rand_core's PCG32 increment constant in the StdRng seed fill, rand_distr's
 * operation order inside log_gamma and the rejection loop, and statrs'
 * erfc_inv.
 */
#include "fg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPSILON_F 1e-6f      /* src/model.rs:10 */
#define MAX_LAMBDA_F 1.0e6f  /* src/model.rs:11 */
#define PI_F 3.14159265358979323846f

/* ------------------------------------------------------------------------- */
/* splitmix64 / mix / mix3 -- src/rng.rs:36-52                                */
/* ------------------------------------------------------------------------- */
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t sm64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t fgo_splitmix64(uint64_t x) { return sm64(x); }
uint64_t fgo_mix(uint64_t seed, uint64_t stream) { return sm64(seed ^ stream); }

/* src/rng.rs:40-44: the add and the rotate come before each splitmix. */
static inline uint64_t mix3(uint64_t seed, uint64_t stream, int64_t a, int64_t b)
{
    uint64_t s = seed ^ stream;
    s = sm64(rotl64(s + (uint64_t)a, 17));
    return sm64(rotl64(s + (uint64_t)b, 41));
}
uint64_t fgo_mix3(uint64_t seed, uint64_t stream, int64_t a, int64_t b)
{
    return mix3(seed, stream, a, b);
}

/* ------------------------------------------------------------------------- */
/* Xoshiro256++ (rand 0.8.5 SmallRng on 64-bit targets)                       */
/* ------------------------------------------------------------------------- */
/* seed_from_u64: four successive splitmix64 outputs, little-endian into the
 * 32-byte seed; an all-zero seed reseeds from 0. */
static inline void xo_seed(uint64_t h, uint64_t s[4])
{
    uint64_t st = h;
    for (int w = 0; w < 4; ++w) {
        s[w] = sm64(st);
        st += 0x9E3779B97F4A7C15ULL;
    }
    if ((s[0] | s[1] | s[2] | s[3]) == 0) xo_seed(0, s);
}

static inline uint64_t xo_next(uint64_t s[4])
{
    const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}
/* next_u32 keeps the upper half (low bits of ++ have linear dependencies). */
static inline uint32_t xo_next_u32(uint64_t s[4]) { return (uint32_t)(xo_next(s) >> 32); }

void fgo_xoshiro_seed(uint64_t h, uint64_t s[4]) { xo_seed(h, s); }
uint64_t fgo_xoshiro_next_u64(uint64_t s[4]) { return xo_next(s); }
uint32_t fgo_xoshiro_next_u32(uint64_t s[4]) { return xo_next_u32(s); }

static inline float f32_from_bits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline uint32_t f32_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }
static inline double f64_from_bits(uint64_t b) { double f; memcpy(&f, &b, 8); return f; }

/* UniformFloat<f32>::new (rand 0.8.5): shrink scale until low + scale*max_rand < high. */
float fgo_uniform_f32_scale(float low, float high)
{
    const float max_rand = f32_from_bits(0x3F800000u | (0xFFFFFFFFu >> 9)) - 1.0f;
    float scale = high - low;
    for (;;) {
        float top = scale * max_rand + low;
        if (!(top >= high)) break;
        scale = f32_from_bits(f32_bits(scale) - 1u);
    }
    return scale;
}

/* UniformFloat<f32>::sample: 23 high bits of next_u32 as a [1,2) mantissa. */
static inline float uni_f32(uint64_t s[4], float low, float scale)
{
    const float v12 = f32_from_bits((xo_next_u32(s) >> 9) | 0x3F800000u);
    const float v01 = v12 - 1.0f;
    return v01 * scale + low;
}
float fgo_uniform_f32_sample(uint64_t s[4], float low, float scale) { return uni_f32(s, low, scale); }

/* Standard f64: 53 high bits times 2^-53, in [0, 1). */
static inline double std_f64_bits(uint64_t v) { return (double)(v >> 11) * (1.0 / 9007199254740992.0); }
/* Open01 f64: 52-bit mantissa in [1,2) minus (1 - 2^-53), in (0, 1). */
static inline double open01_bits(uint64_t v)
{
    return f64_from_bits(0x3FF0000000000000ULL | (v >> 12)) - (1.0 - 2.220446049250313e-16 / 2.0);
}
double fgo_standard_f64(uint64_t s[4]) { return std_f64_bits(xo_next(s)); }
double fgo_open01_f64(uint64_t s[4]) { return open01_bits(xo_next(s)); }

/* ------------------------------------------------------------------------- */
/* rand_distr 0.4.3 log_gamma: Lanczos g=5, six coefficients.                 */
/* ------------------------------------------------------------------------- */
double fgo_log_gamma(double x)
{
    static const double coeff[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                    -1.231739572450155, 0.1208650973866179e-2,
                                    -0.5395239384953e-5};
    const double tmp = x + 5.5;
    const double lg = (x + 0.5) * log(tmp) - tmp;
    double a = 1.000000000190015;
    double denom = x;
    for (int k = 0; k < 6; ++k) {
        denom = denom + 1.0;
        a = a + (coeff[k] / denom);
    }
    return lg + log(2.5066282746310005 * a / x);
}

/* ------------------------------------------------------------------------- */
/* Poisson<f64> (rand_distr 0.4.3): Knuth for lambda < 12, Cauchy rejection   */
/* with the Numerical-Recipes comparison function otherwise.                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    double lambda, exp_lambda, log_lambda, sqrt_2lambda, magic_val;
} poisson_t;

static inline void poisson_new(poisson_t *p, double lambda)
{
    p->lambda = lambda;
    p->log_lambda = log(lambda);
    p->exp_lambda = exp(-lambda);
    p->sqrt_2lambda = sqrt(2.0 * lambda);
    p->magic_val = lambda * p->log_lambda - fgo_log_gamma(1.0 + lambda);
}

static inline double poisson_sample_xo(const poisson_t *p, uint64_t s[4])
{
    if (p->lambda < 12.0) {
        double result = 0.0, prod = 1.0;
        while (prod > p->exp_lambda) {
            prod = prod * std_f64_bits(xo_next(s));
            result = result + 1.0;
        }
        return result - 1.0;
    }
    double result;
    for (;;) {
        double comp_dev;
        for (;;) {
            /* Cauchy(0, 1): median + scale * tan(pi * x) */
            const double x = std_f64_bits(xo_next(s));
            comp_dev = 0.0 + 1.0 * tan(3.14159265358979323846 * x);
            result = p->sqrt_2lambda * comp_dev + p->lambda;
            if (result >= 0.0) break;
        }
        result = floor(result);
        const double check = 0.9 * (1.0 + comp_dev * comp_dev) *
                             exp(result * p->log_lambda - fgo_log_gamma(1.0 + result) - p->magic_val);
        if (std_f64_bits(xo_next(s)) <= check) break;
    }
    return result;
}

double fgo_poisson_sample(uint64_t s[4], double lambda)
{
    poisson_t p;
    poisson_new(&p, lambda);
    return poisson_sample_xo(&p, s);
}

/* Rust `as u32` / `as i32` from floats saturate and map NaN to 0. */
static inline uint32_t sat_u32(double v)
{
    if (!(v > 0.0)) return 0;
    if (v >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)v;
}
static inline int32_t sat_i32f(float v)
{
    if (v != v) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)v;
}
static inline int64_t sat_i64f(float v)
{
    if (v != v) return 0;
    if (v >= 9.2233720368547758e18f) return INT64_MAX;
    if (v <= -9.2233720368547758e18f) return INT64_MIN;
    return (int64_t)v;
}

/* ------------------------------------------------------------------------- */
/* StandardNormal ziggurat (rand_distr 0.4.3).  Tables follow the published   */
/* ZIGNOR recurrence (Doornik 2005) with R = 3.6541528853610088 and           */
/* V = 0.00492867323399, printed with "%.18f" and read back, which is how the */
/* crate's ziggurat_tables.rs literals were produced.  Computed here at run   */
/* time, independently of the product's generated header.                     */
/* ------------------------------------------------------------------------- */
static double ZX[257], ZF[257];
static const double ZIG_R = 3.654152885361008796;
static int zig_ready = 0;

static double zig_round18(double v)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%.18f", v);
    return strtod(buf, NULL);
}

static void zig_init(void)
{
    if (__atomic_load_n(&zig_ready, __ATOMIC_ACQUIRE)) return;
    static double x[257], f[257];
    const double r = 3.6541528853610088, v = 0.00492867323399;
    x[0] = v / exp(-r * r / 2.0);
    x[1] = r;
    for (int i = 2; i < 256; ++i) {
        const double last = x[i - 1];
        x[i] = sqrt(-2.0 * log(v / last + exp(-last * last / 2.0)));
    }
    x[256] = 0.0;
    for (int i = 0; i < 257; ++i) f[i] = exp(-x[i] * x[i] / 2.0);
    for (int i = 0; i < 257; ++i) {
        ZX[i] = zig_round18(x[i]);
        ZF[i] = zig_round18(f[i]);
    }
    __atomic_store_n(&zig_ready, 1, __ATOMIC_RELEASE);
}

void fgo_zig_tables(double x_out[257], double f_out[257])
{
    zig_init();
    memcpy(x_out, ZX, sizeof ZX);
    memcpy(f_out, ZF, sizeof ZF);
}

/* The ziggurat body is generic over the RNG in the crate; it is instantiated
 * here for Xoshiro256++ (radii) and for ChaCha12 (offsets). */
#define ZIGGURAT_BODY(NEXT_U64)                                                          \
    for (;;) {                                                                           \
        const uint64_t bits = NEXT_U64;                                                  \
        const int i = (int)(bits & 0xff);                                                \
        const double u = f64_from_bits(((uint64_t)(1 + 1023) << 52) | (bits >> 12)) - 3.0; \
        const double x = u * ZX[i];                                                      \
        if (fabs(x) < ZX[i + 1]) return x;                                               \
        if (i == 0) {                                                                    \
            double xx = 1.0, yy = 0.0;                                                   \
            while (-2.0 * yy < xx * xx) {                                                \
                const double x_ = open01_bits(NEXT_U64);                                 \
                const double y_ = open01_bits(NEXT_U64);                                 \
                xx = log(x_) / ZIG_R;                                                    \
                yy = log(y_);                                                            \
            }                                                                            \
            return u < 0.0 ? xx - ZIG_R : ZIG_R - xx;                                    \
        }                                                                                \
        if (ZF[i + 1] + (ZF[i] - ZF[i + 1]) * std_f64_bits(NEXT_U64) < exp(-x * x / 2.0)) \
            return x;                                                                    \
    }

static inline double std_normal_xo(uint64_t s[4]) { ZIGGURAT_BODY(xo_next(s)) }

double fgo_standard_normal(uint64_t s[4])
{
    zig_init();
    return std_normal_xo(s);
}

/* LogNormal(mu, sigma).sample = exp(Normal(mu, sigma).sample) */
static inline double lognormal_xo(uint64_t s[4], double mu, double sigma)
{
    return exp(mu + sigma * std_normal_xo(s));
}
double fgo_lognormal_sample(uint64_t s[4], double mu, double sigma)
{
    zig_init();
    return lognormal_xo(s, mu, sigma);
}

/* ------------------------------------------------------------------------- */
/* StdRng = ChaCha12 (rand_chacha 0.3.1) seeded by rand_core 0.6.4's PCG32   */
/* fill; BlockRng over a 64-word buffer (4 blocks).                           */
/* ------------------------------------------------------------------------- */
static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
#define QR(a, b, c, d)                 \
    a += b; d ^= a; d = rotl32(d, 16); \
    c += d; b ^= c; b = rotl32(b, 12); \
    a += b; d ^= a; d = rotl32(d, 8);  \
    c += d; b ^= c; b = rotl32(b, 7);

void fgo_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream, int rounds,
                      uint32_t out[16])
{
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                       key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                       (uint32_t)counter, (uint32_t)(counter >> 32),
                       (uint32_t)stream, (uint32_t)(stream >> 32)};
    uint32_t x[16];
    memcpy(x, in, sizeof x);
    for (int r = 0; r < rounds; r += 2) {
        QR(x[0], x[4], x[8], x[12]) QR(x[1], x[5], x[9], x[13])
        QR(x[2], x[6], x[10], x[14]) QR(x[3], x[7], x[11], x[15])
        QR(x[0], x[5], x[10], x[15]) QR(x[1], x[6], x[11], x[12])
        QR(x[2], x[7], x[8], x[13]) QR(x[3], x[4], x[9], x[14])
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

typedef struct {
    uint32_t key[8];
    uint64_t counter;
    uint32_t buf[64];
    int index;
} stdrng_t;

/* PCG32 (O'Neill's XSH-RR output of a 64-bit LCG) as rand_core 0.6.4's
 * seed_from_u64 runs it: the state is advanced first, then the new state is
 * output.  The multiplier and the output function are pinned by the published
 * pcg32 demo vector; the increment 11634580027462260723 is rand_core's
 * constant and is parity unpinned (no published vector uses it). */
static inline uint32_t pcg32_next(uint64_t *state, uint64_t inc)
{
    const uint64_t s = *state * 6364136223846793005ULL + inc;
    *state = s;
    const uint32_t xorshifted = (uint32_t)(((s >> 18) ^ s) >> 27);
    const uint32_t rot = (uint32_t)(s >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}
uint32_t fgo_pcg32_next(uint64_t *state, uint64_t inc) { return pcg32_next(state, inc); }

static void stdrng_seed_from_u64(stdrng_t *r, uint64_t state)
{
    for (int w = 0; w < 8; ++w) r->key[w] = pcg32_next(&state, 11634580027462260723ULL);
    r->counter = 0;
    r->index = 64;
}

static void stdrng_generate(stdrng_t *r)
{
    for (int b = 0; b < 4; ++b) fgo_chacha_block(r->key, r->counter + (uint64_t)b, 0, 12, r->buf + 16 * b);
    r->counter += 4;
}

static uint64_t stdrng_next_u64(stdrng_t *r)
{
    const int len = 64;
    if (r->index < len - 1) {
        const uint64_t v = ((uint64_t)r->buf[r->index + 1] << 32) | r->buf[r->index];
        r->index += 2;
        return v;
    } else if (r->index >= len) {
        stdrng_generate(r);
        r->index = 2;
        return ((uint64_t)r->buf[1] << 32) | r->buf[0];
    } else {
        const uint64_t lo = r->buf[len - 1];
        stdrng_generate(r);
        r->index = 1;
        return ((uint64_t)r->buf[0] << 32) | lo;
    }
}

void fgo_stdrng_u64(uint64_t seed, uint32_t n, uint64_t *out)
{
    stdrng_t r;
    stdrng_seed_from_u64(&r, seed);
    for (uint32_t k = 0; k < n; ++k) out[k] = stdrng_next_u64(&r);
}

static double std_normal_chacha(stdrng_t *r) { ZIGGURAT_BODY(stdrng_next_u64(r)) }

/* src/rng.rs:9-24 */
int fgo_make_offsets(uint64_t seed, uint32_t n, float sigma_px, float *out_xy)
{
    if (n == 0) return 0;
    zig_init();
    const float feps = 1.1920928955078125e-07f; /* f32::EPSILON */
    const double sigma = (double)(sigma_px > feps ? sigma_px : feps);
    stdrng_t r;
    stdrng_seed_from_u64(&r, fgo_mix(seed, FGO_OFFSET_STREAM));
    for (uint32_t k = 0; k < n; ++k) {
        const double ox = 0.0 + sigma * std_normal_chacha(&r);
        const double oy = 0.0 + sigma * std_normal_chacha(&r);
        out_xy[2 * k] = (float)ox;
        out_xy[2 * k + 1] = (float)oy;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Inverse standard normal CDF (stands in for statrs Normal::inverse_cdf).    */
/* Acklam's rational approximation refined by two Halley steps on erfc.       */
/* ------------------------------------------------------------------------- */
double fgo_inverse_normal_cdf(double p)
{
    static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02,
                                -2.759285104469687e+02, 1.383577518672690e+02,
                                -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02,
                                -1.556989798598866e+02, 6.680131188771972e+01,
                                -1.328068155288572e+01};
    static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01,
                                -2.400758277161838e+00, -2.549732539343734e+00,
                                4.374664141464968e+00, 2.938163982698783e+00};
    static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01,
                                2.445134137142996e+00, 3.754408661907416e+00};
    if (!(p > 0.0)) return -INFINITY;
    if (!(p < 1.0)) return INFINITY;
    const double plow = 0.02425, phigh = 1.0 - plow;
    double x;
    if (p < plow) {
        const double q = sqrt(-2.0 * log(p));
        x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
            ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    } else if (p <= phigh) {
        const double q = p - 0.5, r = q * q;
        x = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
            (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
    } else {
        const double q = sqrt(-2.0 * log(1.0 - p));
        x = -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
            ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    }
    for (int it = 0; it < 2; ++it) {
        /* residual taken on the tail that keeps precision */
        double e;
        if (x > 0.0) e = (1.0 - p) - 0.5 * erfc(x / sqrt(2.0));
        else e = 0.5 * erfc(-x / sqrt(2.0)) - p;
        const double u = e * sqrt(2.0 * 3.14159265358979323846) * exp(x * x / 2.0);
        x = x - u / (1.0 + x * u / 2.0);
    }
    return x;
}

/* ------------------------------------------------------------------------- */
/* Parameters -- src/params.rs:223-261, src/model.rs:181-226,267-298          */
/* ------------------------------------------------------------------------- */
float fgo_default_cell_delta(float radius_mean)
{
    if (radius_mean <= 0.0f) return 1.0f;
    float inv = ceilf(1.0f / radius_mean);
    if (!(inv >= 1.0f)) inv = 1.0f; /* f32::max(1.0) */
    return 1.0f / inv;
}

int fgo_derive(const fgo_params *p, fgo_derived *d)
{
    memset(d, 0, sizeof *d);
    if (p->in_w == 0 || p->in_h == 0) return -1;
    if (!(p->radius_mean > 0.0f) || !(p->zoom > 0.0f)) return -1;
    /* derive_radius_parameters */
    float stddev = 0.0f, log_mu = 0.0f, log_sigma = 0.0f;
    if (p->dist_kind == 1) {
        if (p->radius_stddev == 0.0f) {
            log_mu = logf(p->radius_mean);
            log_sigma = 0.0f;
            stddev = 0.0f;
        } else {
            const float variance_ratio = (p->radius_stddev * p->radius_stddev) /
                                         (p->radius_mean * p->radius_mean);
            const float sigma_sq = logf(1.0f + variance_ratio);
            log_sigma = sqrtf(sigma_sq);
            log_mu = logf(p->radius_mean) - 0.5f * sigma_sq;
            stddev = p->radius_stddev;
        }
    }
    d->radius_stddev = stddev;
    d->log_mu = log_mu;
    d->log_sigma = log_sigma;
    /* resolve_output_size */
    d->in_w = p->in_w;
    d->in_h = p->in_h;
    if (p->size_w > 0) {
        d->out_w = (uint32_t)p->size_w;
        if (p->size_h > 0) d->out_h = (uint32_t)p->size_h;
        else {
            const float aspect = (float)p->in_h / (float)p->in_w;
            float h = roundf((float)p->size_w * aspect);
            if (!(h >= 1.0f)) h = 1.0f;
            d->out_h = (uint32_t)h;
        }
    } else {
        float w = ceilf((float)p->in_w * p->zoom), h = ceilf((float)p->in_h * p->zoom);
        if (!(w >= 1.0f)) w = 1.0f;
        if (!(h >= 1.0f)) h = 1.0f;
        d->out_w = (uint32_t)w;
        d->out_h = (uint32_t)h;
    }
    /* derive_common */
    const float mean_sq = p->radius_mean * p->radius_mean;
    const float variance = stddev * stddev;
    float ms = mean_sq + variance;
    if (!(ms >= EPSILON_F)) ms = EPSILON_F;
    d->inv_e_pi_r2 = 1.0f / (PI_F * ms);
    float rm;
    if (p->max_radius_is_abs) {
        rm = p->max_radius_value;
    } else if (p->dist_kind == 0) {
        rm = p->radius_mean;
    } else {
        const double mu = (double)log_mu, sigma = (double)log_sigma;
        if (sigma == 0.0) rm = (float)exp(mu);
        else rm = (float)exp(mu + sigma * fgo_inverse_normal_cdf((double)p->max_radius_value));
    }
    if (!(rm >= EPSILON_F)) rm = EPSILON_F;
    d->rm = rm;
    float delta = p->cell_delta > 0.0f ? p->cell_delta : fgo_default_cell_delta(p->radius_mean);
    if (!(delta >= EPSILON_F)) delta = EPSILON_F;
    d->delta = delta;
    return 0;
}

/* src/choose.rs:4-26; returns 1 (grain) or 2 (pixel). */
int fgo_choose_algorithm(float radius_mean, float radius_stddev, float rm, uint32_t n_samples,
                         int algo)
{
    if (algo != 0) return algo;
    const float mean = radius_mean > 1e-6f ? radius_mean : 1e-6f;
    const float sigma_ratio = mean > 0.0f ? radius_stddev / mean : 0.0f;
    const float rm_ratio = rm / mean;
    if (sigma_ratio < 0.1f && mean < 0.5f && n_samples <= 64) return 2;
    if (rm_ratio > 8.0f || sigma_ratio > 0.6f || n_samples > 96) return 1;
    if (n_samples <= 24 && rm_ratio < 5.0f) return 2;
    return 1;
}

/* src/model.rs:228-250; returns the max value found. */
float fgo_normalize_plane(const float *src, uint64_t n, float *dst)
{
    float max_value = 0.0f;
    for (uint64_t k = 0; k < n; ++k) {
        const float v = src[k];
        if (v > max_value) max_value = v; /* f32::max from 0.0 never holds NaN */
    }
    const int already = max_value <= 1.0f + EPSILON_F;
    for (uint64_t k = 0; k < n; ++k) {
        float v = src[k];
        if (!already) {
            float denom = max_value + EPSILON_F;
            if (!(denom >= EPSILON_F)) denom = EPSILON_F;
            v = v / denom;
        }
        /* f32::clamp keeps NaN */
        if (v < 0.0f) v = 0.0f;
        if (v > 1.0f - EPSILON_F) v = 1.0f - EPSILON_F;
        dst[k] = v;
    }
    return max_value;
}

/* src/model.rs:252-265 */
void fgo_lambda_plane(const float *normalized, uint64_t n, float inv_e_pi_r2, float *dst)
{
    for (uint64_t k = 0; k < n; ++k) {
        float c = normalized[k];
        if (c < 0.0f) c = 0.0f;
        if (c > 1.0f - EPSILON_F) c = 1.0f - EPSILON_F;
        float safe = 1.0f - c;
        if (!(safe >= EPSILON_F)) safe = EPSILON_F; /* f32::max ignores NaN */
        const float activity = -inv_e_pi_r2 * logf(safe);
        dst[k] = activity < MAX_LAMBDA_F ? activity : MAX_LAMBDA_F; /* f32::min ignores NaN */
    }
}

/* ------------------------------------------------------------------------- */
/* Pixel-wise integrator -- src/pixelwise.rs                                   */
/* ------------------------------------------------------------------------- */
typedef struct {
    const fgo_plane *p;
    const float *lambda;
    float uniform_scale; /* Uniform::new(0, delta) */
    double log_mu, log_sigma;
} pix_ctx;

static inline float lambda_get_clamped(const fgo_plane *p, const float *lambda, int64_t x, int64_t y)
{
    if (p->in_w == 0 || p->in_h == 0) return 0.0f;
    const int64_t w = (int64_t)p->in_w - 1, h = (int64_t)p->in_h - 1;
    const int64_t xi = x < 0 ? 0 : (x > w ? w : x);
    const int64_t yi = y < 0 ? 0 : (y > h ? h : y);
    return lambda[(uint64_t)yi * p->in_w + (uint64_t)xi];
}

static inline float radius_sample(const pix_ctx *c, uint64_t s[4])
{
    if (c->p->dist_kind == 0) return c->p->radius_mean;
    return (float)lognormal_xo(s, c->log_mu, c->log_sigma);
}

/* src/pixelwise.rs:47-106 */
static float evaluate_indicator(const pix_ctx *c, float xg, float yg, fgo_counters *cnt)
{
    const fgo_plane *p = c->p;
    const float rm = p->rm, delta = p->delta;
    if (rm <= 0.0f) return 0.0f;
    const int32_t i0 = sat_i32f(floorf((xg - rm) / delta));
    const int32_t i1 = sat_i32f(floorf((xg + rm) / delta));
    const int32_t j0 = sat_i32f(floorf((yg - rm) / delta));
    const int32_t j1 = sat_i32f(floorf((yg + rm) / delta));
    if (i0 > i1 || j0 > j1) return 0.0f;
    float result = 0.0f;
    int hit = 0;
    for (int64_t id = i0; id <= i1; ++id) {
        for (int64_t jd = j0; jd <= j1; ++jd) {
            uint64_t s[4];
            xo_seed(mix3(p->seed, FGO_CELL_STREAM, (int64_t)(int32_t)id, (int64_t)(int32_t)jd), s);
            const float sample_x = (float)(int32_t)id * delta;
            const float sample_y = (float)(int32_t)jd * delta;
            const int64_t ix = sat_i64f(floorf(sample_x));
            const int64_t iy = sat_i64f(floorf(sample_y));
            const float lambda_cell = lambda_get_clamped(p, c->lambda, ix, iy);
            if (lambda_cell <= 0.0f) continue;
            const float expected = lambda_cell * delta * delta;
            if (expected <= 0.0f) continue;
            if (cnt) cnt->cell_visits++;
            poisson_t po;
            poisson_new(&po, (double)expected);
            const uint32_t q = sat_u32(poisson_sample_xo(&po, s));
            for (uint32_t g = 0; g < q; ++g) {
                const float cx = sample_x + uni_f32(s, 0.0f, c->uniform_scale);
                const float cy = sample_y + uni_f32(s, 0.0f, c->uniform_scale);
                float radius = radius_sample(c, s);
                if (radius > p->rm) radius = p->rm;
                if (radius <= 0.0f) continue;
                const float dx = xg - cx;
                const float dy = yg - cy;
                const int inside = dx * dx + dy * dy <= radius * radius;
                if (cnt) {
                    cnt->tests_full++;
                    if (!hit) cnt->tests++;
                }
                if (inside && !hit) {
                    hit = 1;
                    result = 1.0f;
                    if (!cnt) return result; /* early exit, as the reference does */
                }
            }
        }
    }
    return result;
}

int fgo_render_pixelwise(const fgo_plane *p, const float *lambda, const float *offsets_input_xy,
                         uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, float *out,
                         int threads, fgo_counters *counters)
{
    if (!p || !lambda || !out || x1 > p->out_w || y1 > p->out_h || x0 > x1 || y0 > y1) return -1;
    if (p->n_samples > 0 && !offsets_input_xy) return -1;
    zig_init();
    pix_ctx c;
    c.p = p;
    c.lambda = lambda;
    c.uniform_scale = fgo_uniform_f32_scale(0.0f, p->delta);
    c.log_mu = (double)p->log_mu;
    c.log_sigma = (double)p->log_sigma;
    const uint32_t nsamp = p->n_samples > 1 ? p->n_samples : 1;
    const float inv_samples = 1.0f / (float)nsamp;
    const float inv_zoom = 1.0f / p->zoom;
    const uint32_t w = x1 - x0;
    fgo_counters total = {0, 0, 0, 0, 0};
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int64_t yy = (int64_t)y0; yy < (int64_t)y1; ++yy) {
        fgo_counters local = {0, 0, 0, 0, 0};
        for (uint32_t x = x0; x < x1; ++x) {
            float sum = 0.0f;
            for (uint32_t k = 0; k < p->n_samples; ++k) {
                const float xg = ((float)x + 0.5f) * inv_zoom - offsets_input_xy[2 * k];
                const float yg = ((float)yy + 0.5f) * inv_zoom - offsets_input_xy[2 * k + 1];
                const float v = evaluate_indicator(&c, xg, yg, counters ? &local : NULL);
                sum += v;
                if (counters) {
                    local.points++;
                    if (v != 0.0f) local.hits++;
                }
            }
            out[(uint64_t)(yy - y0) * w + (x - x0)] = sum * inv_samples;
        }
        if (counters) {
#ifdef _OPENMP
#pragma omp critical
#endif
            {
                total.points += local.points;
                total.cell_visits += local.cell_visits;
                total.tests += local.tests;
                total.tests_full += local.tests_full;
                total.hits += local.hits;
            }
        }
    }
    if (counters) *counters = total;
    return 0;
}

int fgo_cell_dump(const fgo_plane *p, const float *lambda, int32_t i, int32_t j, uint64_t *hash,
                  uint64_t state[4], float *grains_xyr, int max_grains)
{
    zig_init();
    const float delta = p->delta;
    const float scale = fgo_uniform_f32_scale(0.0f, delta);
    pix_ctx c;
    c.p = p;
    c.lambda = lambda;
    c.uniform_scale = scale;
    c.log_mu = (double)p->log_mu;
    c.log_sigma = (double)p->log_sigma;
    uint64_t s[4];
    const uint64_t h = mix3(p->seed, FGO_CELL_STREAM, (int64_t)i, (int64_t)j);
    xo_seed(h, s);
    if (hash) *hash = h;
    if (state) memcpy(state, s, sizeof s);
    const float sample_x = (float)i * delta, sample_y = (float)j * delta;
    const float lambda_cell = lambda_get_clamped(p, lambda, sat_i64f(floorf(sample_x)), sat_i64f(floorf(sample_y)));
    if (lambda_cell <= 0.0f) return 0;
    const float expected = lambda_cell * delta * delta;
    if (expected <= 0.0f) return 0;
    poisson_t po;
    poisson_new(&po, (double)expected);
    const uint32_t q = sat_u32(poisson_sample_xo(&po, s));
    for (uint32_t g = 0; g < q; ++g) {
        const float cx = sample_x + uni_f32(s, 0.0f, scale);
        const float cy = sample_y + uni_f32(s, 0.0f, scale);
        float radius = radius_sample(&c, s);
        if (radius > p->rm) radius = p->rm;
        if ((int)g < max_grains && grains_xyr) {
            grains_xyr[3 * g] = cx;
            grains_xyr[3 * g + 1] = cy;
            grains_xyr[3 * g + 2] = radius;
        }
    }
    return (int)q;
}

/* The Poisson count of cell (i, j) alone (src/pixelwise.rs:66-82: cell_rng,
 * the clamped lambda lookup, Poisson::new(lambda * delta * delta), `as u32`). */
uint32_t fgo_cell_count(const fgo_plane *p, const float *lambda, int32_t i, int32_t j)
{
    const float delta = p->delta;
    uint64_t s[4];
    xo_seed(mix3(p->seed, FGO_CELL_STREAM, (int64_t)i, (int64_t)j), s);
    const float sample_x = (float)i * delta, sample_y = (float)j * delta;
    const float lambda_cell = lambda_get_clamped(p, lambda, sat_i64f(floorf(sample_x)), sat_i64f(floorf(sample_y)));
    if (lambda_cell <= 0.0f) return 0;
    const float expected = lambda_cell * delta * delta;
    if (expected <= 0.0f) return 0;
    poisson_t po;
    poisson_new(&po, (double)expected);
    return sat_u32(poisson_sample_xo(&po, s));
}

/* ------------------------------------------------------------------------- */
/* Grain-wise integrator -- src/grainwise.rs                                   */
/* ------------------------------------------------------------------------- */
/* src/grainwise.rs:126-142 */
static inline int bounds(float center, float radius, int32_t limit, int32_t *lo, int32_t *hi)
{
    int32_t mn = sat_i32f(ceilf((center - radius) - 0.5f));
    int32_t mx = sat_i32f(floorf((center + radius) - 0.5f));
    if (mx < mn) return 0;
    if (limit <= 0) return 0;
    const int32_t last = limit - 1;
    if (mn > last || mx < 0) return 0;
    mn = mn < 0 ? 0 : (mn > last ? last : mn);
    mx = mx < 0 ? 0 : (mx > last ? last : mx);
    if (mn > mx) return 0;
    *lo = mn;
    *hi = mx;
    return 1;
}

int fgo_render_grainwise(const fgo_plane *p, const float *lambda, const float *offsets_xy,
                         float *out, uint64_t *bitsets_out, int threads, uint64_t *pairs_out)
{
    fgo_grain_counters c;
    const int rc = fgo_render_grainwise_counted(p, lambda, offsets_xy, out, bitsets_out, threads, &c);
    if (rc == 0 && pairs_out) *pairs_out = c.pairs;
    return rc;
}

int fgo_render_grainwise_counted(const fgo_plane *p, const float *lambda, const float *offsets_xy,
                                 float *out, uint64_t *bitsets_out, int threads,
                                 fgo_grain_counters *counters)
{
    if (!p || !lambda || !out) return -1;
    if (p->n_samples > 0 && !offsets_xy) return -1;
    zig_init();
    const uint64_t total = (uint64_t)p->out_w * p->out_h;
    const uint32_t lanes = (p->n_samples + 63) / 64;
    uint64_t *bits = bitsets_out;
    int own = 0;
    if (!bits) {
        bits = (uint64_t *)calloc(total * (lanes ? lanes : 1), sizeof(uint64_t));
        if (!bits) return -2;
        own = 1;
    } else {
        memset(bits, 0, total * lanes * sizeof(uint64_t));
    }
    pix_ctx c;
    c.p = p;
    c.lambda = lambda;
    c.uniform_scale = fgo_uniform_f32_scale(0.0f, 1.0f);
    c.log_mu = (double)p->log_mu;
    c.log_sigma = (double)p->log_sigma;
    const float zoom = p->zoom;
    const uint32_t nsamp = p->n_samples > 1 ? p->n_samples : 1;
    const float inv_samples = 1.0f / (float)nsamp;
    uint64_t grains = 0, pairs = 0, tests = 0, hits = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(+ : grains, pairs, tests, hits)
#endif
    for (int64_t y = 0; y < (int64_t)p->in_h; ++y) {
        for (uint32_t x = 0; x < p->in_w; ++x) {
            const float lambda_val = lambda[(uint64_t)y * p->in_w + x];
            if (lambda_val <= 0.0f) continue;
            uint64_t s[4];
            xo_seed(mix3(p->seed, FGO_PIXEL_STREAM, (int64_t)(int32_t)x, (int64_t)(int32_t)y), s);
            poisson_t po;
            poisson_new(&po, (double)lambda_val);
            const uint32_t q = sat_u32(poisson_sample_xo(&po, s));
            for (uint32_t g = 0; g < q; ++g) {
                const float cx = (float)x + uni_f32(s, 0.0f, c.uniform_scale);
                const float cy = (float)y + uni_f32(s, 0.0f, c.uniform_scale);
                float radius = radius_sample(&c, s);
                if (radius > p->rm) radius = p->rm;
                if (radius <= 0.0f) continue;
                const float radius_out = radius * zoom;
                if (radius_out <= 0.0f) continue;
                const float radius_sq = radius_out * radius_out;
                grains++;
                for (uint32_t k = 0; k < p->n_samples; ++k) {
                    pairs++;
                    const float tx = (cx * zoom) + offsets_xy[2 * k];
                    const float ty = (cy * zoom) + offsets_xy[2 * k + 1];
                    int32_t xmn, xmx, ymn, ymx;
                    if (!bounds(tx, radius_out, (int32_t)p->out_w, &xmn, &xmx)) continue;
                    if (!bounds(ty, radius_out, (int32_t)p->out_h, &ymn, &ymx)) continue;
                    const uint32_t lane_idx = k / 64;
                    const uint64_t mask = 1ULL << (k % 64);
                    for (int32_t oy = ymn; oy <= ymx; ++oy) {
                        const float center_y = (float)oy + 0.5f;
                        const float dy = center_y - ty;
                        const float dy_sq = dy * dy;
                        if (dy_sq > radius_sq) continue;
                        for (int32_t ox = xmn; ox <= xmx; ++ox) {
                            const float center_x = (float)ox + 0.5f;
                            const float dx = center_x - tx;
                            tests++;
                            if (dx * dx + dy_sq <= radius_sq) {
                                const uint64_t idx = (uint64_t)oy * p->out_w + (uint64_t)ox;
                                __atomic_fetch_or(&bits[idx * lanes + lane_idx], mask, __ATOMIC_RELAXED);
                                hits++;
                            }
                        }
                    }
                }
            }
        }
    }
    for (uint64_t px = 0; px < total; ++px) {
        uint32_t count = 0;
        for (uint32_t l = 0; l < lanes; ++l) count += (uint32_t)__builtin_popcountll(bits[px * lanes + l]);
        out[px] = (float)count * inv_samples;
    }
    if (own) free(bits);
    if (counters) {
        counters->grains = grains;
        counters->pairs = pairs;
        counters->tests = tests;
        counters->hits = hits;
    }
    return 0;
}

int fgo_pixel_dump(const fgo_plane *p, const float *lambda, int32_t x, int32_t y, uint64_t *hash,
                   uint64_t state[4], float *grains_xyr, int max_grains)
{
    zig_init();
    pix_ctx c;
    c.p = p;
    c.lambda = lambda;
    c.uniform_scale = fgo_uniform_f32_scale(0.0f, 1.0f);
    c.log_mu = (double)p->log_mu;
    c.log_sigma = (double)p->log_sigma;
    uint64_t s[4];
    const uint64_t h = mix3(p->seed, FGO_PIXEL_STREAM, (int64_t)x, (int64_t)y);
    xo_seed(h, s);
    if (hash) *hash = h;
    if (state) memcpy(state, s, sizeof s);
    const float lambda_val = lambda[(uint64_t)y * p->in_w + (uint64_t)x];
    if (lambda_val <= 0.0f) return 0;
    poisson_t po;
    poisson_new(&po, (double)lambda_val);
    const uint32_t q = sat_u32(poisson_sample_xo(&po, s));
    for (uint32_t g = 0; g < q; ++g) {
        const float cx = (float)x + uni_f32(s, 0.0f, c.uniform_scale);
        const float cy = (float)y + uni_f32(s, 0.0f, c.uniform_scale);
        float radius = radius_sample(&c, s);
        if (radius > p->rm) radius = p->rm;
        if ((int)g < max_grains && grains_xyr) {
            grains_xyr[3 * g] = cx;
            grains_xyr[3 * g + 1] = cy;
            grains_xyr[3 * g + 2] = radius;
        }
    }
    return (int)q;
}

/* Digest of one grain record at position k of its cell: splitmix64's
 * finaliser of the f32 bits of (cx, cy, r) and k, summed per cell mod 2^64
 * (tests/cellcheck.py computes the same from the device's grain buffer). */
static inline uint64_t grain_digest(float cx, float cy, float r, uint64_t k)
{
    uint64_t v = ((uint64_t)f32_bits(cx) << 32) | (uint64_t)f32_bits(cy);
    v ^= (uint64_t)f32_bits(r) * 0x9E3779B97F4A7C15ULL;
    v ^= k * 0xD1B54A32D192ED03ULL;
    return sm64(v);
}

int fgo_cell_digests(const fgo_plane *p, const float *lambda, int32_t i0, int32_t j0, uint32_t w, uint32_t h,
                     uint32_t *q_out, uint64_t *digest_out, int threads)
{
    if (!p || !lambda || !q_out || !digest_out) return -1;
    zig_init();
    const float delta = p->delta;
    pix_ctx c;
    c.p = p;
    c.lambda = lambda;
    c.uniform_scale = fgo_uniform_f32_scale(0.0f, delta);
    c.log_mu = (double)p->log_mu;
    c.log_sigma = (double)p->log_sigma;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
#endif
    for (int64_t jj = 0; jj < (int64_t)h; ++jj) {
        for (uint32_t ii = 0; ii < w; ++ii) {
            const int32_t i = i0 + (int32_t)ii, j = j0 + (int32_t)jj;
            const uint64_t at = (uint64_t)jj * w + ii;
            uint64_t s[4];
            xo_seed(mix3(p->seed, FGO_CELL_STREAM, (int64_t)i, (int64_t)j), s);
            const float sample_x = (float)i * delta, sample_y = (float)j * delta;
            const float lambda_cell =
                lambda_get_clamped(p, lambda, sat_i64f(floorf(sample_x)), sat_i64f(floorf(sample_y)));
            uint32_t q = 0;
            uint64_t dg = 0;
            const float expected = lambda_cell * delta * delta;
            if (lambda_cell > 0.0f && expected > 0.0f) {
                poisson_t po;
                poisson_new(&po, (double)expected);
                q = sat_u32(poisson_sample_xo(&po, s));
                for (uint32_t g = 0; g < q; ++g) {
                    const float cx = sample_x + uni_f32(s, 0.0f, c.uniform_scale);
                    const float cy = sample_y + uni_f32(s, 0.0f, c.uniform_scale);
                    float radius = radius_sample(&c, s);
                    if (radius > p->rm) radius = p->rm;
                    dg += grain_digest(cx, cy, radius, g);
                }
            }
            q_out[at] = q;
            digest_out[at] = dg;
        }
    }
    return 0;
}

int fgo_pixel_counts(const fgo_plane *p, const float *lambda, uint32_t *q_out, uint64_t *state_out, int threads)
{
    if (!p || !lambda || !q_out) return -1;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
#endif
    for (int64_t y = 0; y < (int64_t)p->in_h; ++y) {
        for (uint32_t x = 0; x < p->in_w; ++x) {
            const uint64_t at = (uint64_t)y * p->in_w + x;
            const float lambda_val = lambda[at];
            uint64_t s[4] = {0, 0, 0, 0};
            uint32_t q = 0;
            if (lambda_val > 0.0f) {
                xo_seed(mix3(p->seed, FGO_PIXEL_STREAM, (int64_t)(int32_t)x, (int64_t)y), s);
                poisson_t po;
                poisson_new(&po, (double)lambda_val);
                q = sat_u32(poisson_sample_xo(&po, s));
            }
            q_out[at] = q;
            if (state_out) memcpy(state_out + 4 * at, s, sizeof s);
        }
    }
    return 0;
}

/* The C library's f64 math (what Rust's f64::exp/ln/tan/sqrt call on Linux,
 * and through them rand_distr), element-wise, for comparison with the
 * device: fn as FG_MATH_* (0 exp(-x), 1 exp(x), 2 log(x), 3 tan(pi x),
 * 4 log_gamma(x), 5 sqrt(x), 6 (double)(float)exp(x)). */
int fgo_glibc_math(int fn, const double *in, int64_t n, double *out, int threads)
{
    if (fn < 0 || fn > 6 || n < 0) return -1;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int64_t i = 0; i < n; ++i) {
        const double x = in[i];
        double r = 0.0;
        switch (fn) {
        case 0: r = exp(-x); break;
        case 1: r = exp(x); break;
        case 2: r = log(x); break;
        case 3: r = tan(3.14159265358979323846 * x); break;
        case 4: r = fgo_log_gamma(x); break;
        case 5: r = sqrt(x); break;
        case 6: r = (double)(float)exp(x); break;
        }
        out[i] = r;
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * Colour workspaces (src/color.rs).  Pins (tests/test_oracle_colour.py):
 * BT.709's defining identities on the constants and closed forms on pure
 * colours (Y of white and grey, Cb of blue and Cr of red = 1/2, grey chroma
 * = 0), the exact inverse over every 8-bit colour (split then merge returns
 * it), to_u8 on every level k/255 and its saturation, and nearest resize as
 * pixel replication at integer zoom and as the identity at equal size.
 * ------------------------------------------------------------------------ */
static const float FGO_YR = 0.2126f, FGO_YG = 0.7152f, FGO_YB = 0.0722f;  /* src/color.rs:9-11 */
static const float FGO_CB_DENOM = 1.8556f, FGO_CR_DENOM = 1.5748f;         /* src/color.rs:12-13 */

/* f32::clamp(0.0, 1.0): NaN passes through (src/color.rs:233-235) */
static float fgo_clamp01(float v)
{
    if (v < 0.0f) return 0.0f;
    if (v > 1.0f) return 1.0f;
    return v;
}

/* src/color.rs:237-239: (clamp01(v) * 255 + 0.5).floor() as u8, where `as`
 * saturates and sends NaN to 0 */
uint8_t fgo_to_u8(float v)
{
    const float f = floorf(fgo_clamp01(v) * 255.0f + 0.5f);
    if (!(f > 0.0f)) return 0;
    if (f >= 255.0f) return 255;
    return (uint8_t)f;
}

/* src/color.rs:158-213: load_rgb_workspace (mode 1) and load_luma_workspace
 * (mode 0) on an interleaved RGB f32 image of n pixels. */
int fgo_color_split(int mode, const float *rgb, uint64_t n, float *p0, float *p1, float *p2)
{
    if (mode != 0 && mode != 1) return -1;
    for (uint64_t k = 0; k < n; ++k) {
        const float r = fgo_clamp01(rgb[3 * k]);
        const float g = fgo_clamp01(rgb[3 * k + 1]);
        const float b = fgo_clamp01(rgb[3 * k + 2]);
        if (mode == 1) {
            p0[k] = r;
            p1[k] = g;
            p2[k] = b;
            continue;
        }
        const float luma = FGO_YR * r + FGO_YG * g + FGO_YB * b;
        p0[k] = fgo_clamp01(luma);
        p1[k] = (b - luma) / FGO_CB_DENOM;
        p2[k] = (r - luma) / FGO_CR_DENOM;
    }
    return 0;
}

/* Plane::resize_nearest (src/model.rs:77-98): source index
 * round(clamp((x + 0.5) * (w / nw) - 0.5, 0, w - 1)), rounding half away from
 * zero as f32::round does. */
void fgo_resize_nearest(const float *src, uint32_t w, uint32_t h, uint32_t nw, uint32_t nh, float *dst)
{
    if (nw == w && nh == h) {
        memcpy(dst, src, (size_t)w * h * sizeof(float));
        return;
    }
    if (nw == 0 || nh == 0 || w == 0 || h == 0) {
        memset(dst, 0, (size_t)nw * nh * sizeof(float));
        return;
    }
    const float sx = (float)w / (float)nw, sy = (float)h / (float)nh;
    for (uint32_t y = 0; y < nh; ++y) {
        float fy = ((float)y + 0.5f) * sy - 0.5f;
        fy = fy < 0.0f ? 0.0f : (fy > (float)(h - 1) ? (float)(h - 1) : fy);
        const size_t syi = (size_t)roundf(fy);
        for (uint32_t x = 0; x < nw; ++x) {
            float fx = ((float)x + 0.5f) * sx - 0.5f;
            fx = fx < 0.0f ? 0.0f : (fx > (float)(w - 1) ? (float)(w - 1) : fx);
            dst[(size_t)y * nw + x] = src[syi * w + (size_t)roundf(fx)];
        }
    }
}

/* Workspace::into_rgb_image (src/color.rs:66-114): mode 0 takes Y at w x h
 * and Cb/Cr at cw x ch (resized nearest to w x h when they differ); mode 1
 * takes R, G, B at w x h.  rgb8 gets w x h interleaved 8-bit RGB. */
int fgo_color_merge(int mode, const float *p0, const float *p1, const float *p2, uint32_t w, uint32_t h,
                    uint32_t cw, uint32_t ch, uint8_t *rgb8)
{
    const size_t n = (size_t)w * h;
    if (mode == 1) {
        for (size_t k = 0; k < n; ++k) {
            rgb8[3 * k] = fgo_to_u8(p0[k]);
            rgb8[3 * k + 1] = fgo_to_u8(p1[k]);
            rgb8[3 * k + 2] = fgo_to_u8(p2[k]);
        }
        return 0;
    }
    if (mode != 0) return -1;
    float *cb = (float *)malloc(n * sizeof(float) + 1);
    float *cr = (float *)malloc(n * sizeof(float) + 1);
    if (!cb || !cr) {
        free(cb);
        free(cr);
        return -1;
    }
    fgo_resize_nearest(p1, cw, ch, w, h, cb);
    fgo_resize_nearest(p2, cw, ch, w, h, cr);
    for (size_t k = 0; k < n; ++k) {
        const float y = p0[k];
        const float r = fgo_clamp01(y + FGO_CR_DENOM * cr[k]);
        const float b = fgo_clamp01(y + FGO_CB_DENOM * cb[k]);
        const float g = fgo_clamp01((y - FGO_YR * r - FGO_YB * b) / FGO_YG);
        rgb8[3 * k] = fgo_to_u8(r);
        rgb8[3 * k + 1] = fgo_to_u8(g);
        rgb8[3 * k + 2] = fgo_to_u8(b);
    }
    free(cb);
    free(cr);
    return 0;
}

/* The BT.709 constants as the reference writes them (for the identity checks). */
void fgo_color_constants(float out[5])
{
    out[0] = FGO_YR;
    out[1] = FGO_YG;
    out[2] = FGO_YB;
    out[3] = FGO_CB_DENOM;
    out[4] = FGO_CR_DENOM;
}

int fgo_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
