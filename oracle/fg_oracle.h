/*
 * fg_oracle.h -- CPU oracle for the film-grain hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference's CPU integrators and of the third-party random-number algorithms
 * they call.  It exists so that tests, the smoke check and the CPU-baseline leg
 * of the benchmark can check the HIP product against it.  Nothing in the
 * product (film_grain_amd/) links, loads or calls it.
 *
 * Reference files restated (paths are relative to the reference tree):
 *   src/rng.rs          splitmix64 / mix / mix3, cell_rng, pixel_rng, make_offsets
 *   src/pixelwise.rs    render_pixelwise, evaluate_indicator
 *   src/grainwise.rs    render_grainwise, bounds
 *   src/model.rs        RadiusProfile, derive_common, normalize_plane, lambda_plane,
 *                       resolve_output_size, Plane::get_clamped
 *   src/params.rs       derive_radius_parameters, default_cell_delta
 *   src/choose.rs       choose_algorithm
 *   src/color.rs        colour split / merge, to_u8 (with model.rs resize_nearest)
 * Third-party algorithms restated (versions pinned by the reference Cargo.lock):
 *   rand 0.8.5        Xoshiro256PlusPlus (SmallRng), seed_from_u64, UniformFloat<f32>,
 *                     Standard f64, Open01 f64
 *   rand_core 0.6.4   SeedableRng::seed_from_u64 (PCG32 fill) used by StdRng
 *   rand_chacha 0.3.1 ChaCha12 block RNG (StdRng), BlockRng::next_u64
 *   rand_distr 0.4.3  Poisson (Knuth + Cauchy rejection), log_gamma, Cauchy,
 *                     StandardNormal ziggurat + tables, Normal, LogNormal
 *   statrs 0.16.1     Normal::inverse_cdf (restated by a refined Acklam inverse;
 *                     parity unpinned beyond the f32 cast that consumes it)
 *
 * Pinning status of each function is stated where it is defined in
 * fg_oracle.c and checked by tests/test_oracle_kats.py.
 */
#ifndef FG_ORACLE_H
#define FG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stream constants (src/rng.rs:5-7) ---- */
#define FGO_OFFSET_STREAM 0x9E3779B97F4A7C15ULL
#define FGO_CELL_STREAM 0xA24B1C30BEBCCF59ULL
#define FGO_PIXEL_STREAM 0x6935FA5C55F65F1BULL

/* ---- primitives ---- */
uint64_t fgo_splitmix64(uint64_t x);
uint64_t fgo_mix(uint64_t seed, uint64_t stream);
uint64_t fgo_mix3(uint64_t seed, uint64_t stream, int64_t a, int64_t b);
void fgo_xoshiro_seed(uint64_t h, uint64_t s[4]);
uint64_t fgo_xoshiro_next_u64(uint64_t s[4]);
uint32_t fgo_xoshiro_next_u32(uint64_t s[4]);
float fgo_uniform_f32_scale(float low, float high);
float fgo_uniform_f32_sample(uint64_t s[4], float low, float scale);
double fgo_standard_f64(uint64_t s[4]);
double fgo_open01_f64(uint64_t s[4]);
double fgo_log_gamma(double x);
double fgo_poisson_sample(uint64_t s[4], double lambda);
double fgo_standard_normal(uint64_t s[4]);
double fgo_lognormal_sample(uint64_t s[4], double mu, double sigma);
void fgo_zig_tables(double x_out[257], double f_out[257]);
void fgo_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream,
                      int rounds, uint32_t out[16]);
uint32_t fgo_pcg32_next(uint64_t *state, uint64_t inc);
void fgo_stdrng_u64(uint64_t seed, uint32_t n, uint64_t *out);
int fgo_make_offsets(uint64_t seed, uint32_t n, float sigma_px, float *out_xy);
double fgo_inverse_normal_cdf(double p);

/* ---- derived parameters (src/params.rs:223-261, src/model.rs:181-226) ---- */
typedef struct {
    float radius_mean;     /* mu_r, input pixels */
    float radius_stddev;   /* sigma_r as given (ignored for const) */
    uint32_t dist_kind;    /* 0 const, 1 lognormal */
    float zoom;
    float sigma_px;
    uint32_t n_samples;
    int32_t max_radius_is_abs; /* 0: quantile, 1: absolute */
    float max_radius_value;    /* quantile p or absolute radius */
    float cell_delta;          /* <= 0: default_cell_delta(radius_mean) */
    uint64_t seed;
    uint32_t in_w, in_h;
    int32_t size_w, size_h;    /* forced output size; <= 0 means unset */
} fgo_params;

typedef struct {
    uint32_t in_w, in_h, out_w, out_h;
    float inv_e_pi_r2, rm, delta;
    float radius_stddev;      /* as stored in Params after derive_radius_parameters */
    float log_mu, log_sigma;  /* f32, exactly as Params holds them */
} fgo_derived;

int fgo_derive(const fgo_params *p, fgo_derived *d);
float fgo_default_cell_delta(float radius_mean);
int fgo_choose_algorithm(float radius_mean, float radius_stddev, float rm,
                         uint32_t n_samples, int algo); /* 0 auto,1 grain,2 pixel */

/* ---- planes (src/model.rs:228-265) ---- */
float fgo_normalize_plane(const float *src, uint64_t n, float *dst);
void fgo_lambda_plane(const float *normalized, uint64_t n, float inv_e_pi_r2, float *dst);

/* ---- integrators ---- */
typedef struct {
    uint32_t in_w, in_h, out_w, out_h, n_samples;
    uint32_t dist_kind;   /* 0 const, 1 lognormal */
    uint64_t seed;
    float zoom, delta, rm, radius_mean;
    float log_mu, log_sigma;   /* widened to f64 inside, as RadiusProfile::new does */
} fgo_plane;

typedef struct {
    uint64_t points;      /* (pixel, sample) pairs evaluated */
    uint64_t cell_visits; /* cells visited with lambda > 0 */
    uint64_t tests;       /* disk tests in reference order with early exit (T) */
    uint64_t tests_full;  /* disk tests without early exit (T_full) */
    uint64_t hits;        /* covered pairs */
} fgo_counters;

/* Renders output pixels [x0,x1) x [y0,y1) of the pixel-wise integrator into
 * out[(y-y0)*(x1-x0) + (x-x0)].  threads <= 0 uses all OpenMP threads.
 * counters may be NULL. Returns 0, or -1 on invalid arguments. */
int fgo_render_pixelwise(const fgo_plane *p, const float *lambda,
                         const float *offsets_input_xy, uint32_t x0, uint32_t y0,
                         uint32_t x1, uint32_t y1, float *out, int threads,
                         fgo_counters *counters);

/* Full-frame grain-wise integrator; bitsets_out (optional, out_w*out_h*lanes
 * u64 words with lanes = ceil(N/64)) receives the sample bitsets. */
int fgo_render_grainwise(const fgo_plane *p, const float *lambda,
                         const float *offsets_xy, float *out, uint64_t *bitsets_out,
                         int threads, uint64_t *pairs_out);

/* The same with the work counters: grains drawn with a positive output
 * radius, (grain, sample) pairs (grains * N), disk tests (pixels of each
 * pair's bounds() box in rows whose dy^2 <= r^2, src/grainwise.rs:82-100) and
 * the tests that pass. */
typedef struct {
    uint64_t grains, pairs, tests, hits;
} fgo_grain_counters;
int fgo_render_grainwise_counted(const fgo_plane *p, const float *lambda,
                                 const float *offsets_xy, float *out, uint64_t *bitsets_out,
                                 int threads, fgo_grain_counters *counters);

/* Per-cell dump of the pixel-wise generator: hash, seeded xoshiro state,
 * Poisson count and up to max_grains grains (cx, cy, r) for cell (i, j).
 * Returns q (the full count) or -1 on error. */
int fgo_cell_dump(const fgo_plane *p, const float *lambda, int32_t i, int32_t j,
                  uint64_t *hash, uint64_t state[4], float *grains_xyr,
                  int max_grains);

/* Poisson count of cell (i, j) only (no grain draws). */
uint32_t fgo_cell_count(const fgo_plane *p, const float *lambda, int32_t i, int32_t j);

/* Per-input-pixel dump of the grain-wise generator. */
int fgo_pixel_dump(const fgo_plane *p, const float *lambda, int32_t x, int32_t y,
                   uint64_t *hash, uint64_t state[4], float *grains_xyr,
                   int max_grains);

/* Every cell of the w x h grid from (i0, j0): the Poisson count and a digest
 * of its grains (the sum mod 2^64 of splitmix64's finaliser of each grain's
 * f32 bits and position), row-major.  And every input pixel's grain-wise
 * Poisson count, with the stream state after the count draws (state_out,
 * optional, 4 words per pixel; zero where lambda <= 0). */
int fgo_cell_digests(const fgo_plane *p, const float *lambda, int32_t i0, int32_t j0, uint32_t w, uint32_t h,
                     uint32_t *q_out, uint64_t *digest_out, int threads);
int fgo_pixel_counts(const fgo_plane *p, const float *lambda, uint32_t *q_out, uint64_t *state_out, int threads);

/* The C library's f64 math element-wise (fn: 0 exp(-x), 1 exp(x), 2 log(x),
 * 3 tan(pi x), 4 log_gamma(x), 5 sqrt(x), 6 (double)(float)exp(x)). */
int fgo_glibc_math(int fn, const double *in, int64_t n, double *out, int threads);

/* ---- colour workspaces (src/color.rs, src/model.rs:77-98) ---- */
uint8_t fgo_to_u8(float v);
int fgo_color_split(int mode, const float *rgb, uint64_t n, float *p0, float *p1, float *p2); /* 0 luma, 1 rgb */
void fgo_resize_nearest(const float *src, uint32_t w, uint32_t h, uint32_t nw, uint32_t nh, float *dst);
int fgo_color_merge(int mode, const float *p0, const float *p1, const float *p2, uint32_t w, uint32_t h,
                    uint32_t cw, uint32_t ch, uint8_t *rgb8);
void fgo_color_constants(float out[5]); /* Kr, Kg, Kb, Cb denominator, Cr denominator */

int fgo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
