// fg_model.cpp -- planes, radius profile, derived parameters, offsets and
// routing (src/model.rs, src/rng.rs:9-24, src/choose.rs, src/lib.rs:205-222).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "../fg_rng.h"
#include "fg_host.h"

namespace film_grain {

namespace {
constexpr float kEpsilon = 1e-6f;      // src/model.rs:10
constexpr float kMaxLambda = 1.0e6f;   // src/model.rs:11
constexpr float kPiF = 3.14159265358979323846f;
const double kZX[257] = {FG_ZIG_NORM_X_VALUES};
const double kZF[257] = {FG_ZIG_NORM_F_VALUES};

inline float max_nan(float a, float b)  // f32::max: a NaN operand yields the other
{
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}

// rand 0.8.5 StdRng: ChaCha with 12 rounds (rand_chacha 0.3.1), 64-bit block
// counter, zero stream, a 64-word buffer refilled four blocks at a time, seeded
// by rand_core 0.6.4's SeedableRng::seed_from_u64 (PCG32 output words as key).
class StdRng {
public:
    explicit StdRng(uint64_t state)
    {
        for (int w = 0; w < 8; ++w) {
            state = state * 6364136223846793005ULL + 11634580027462260723ULL;
            const uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
            const uint32_t rot = (uint32_t)(state >> 59);
            key_[w] = (xs >> rot) | (xs << ((32u - rot) & 31u));
        }
    }

    uint64_t next_u64()
    {
        if (index_ < 63) {
            const uint64_t v = (uint64_t)buf_[index_] | ((uint64_t)buf_[index_ + 1] << 32);
            index_ += 2;
            return v;
        }
        if (index_ >= 64) {
            refill();
            index_ = 2;
            return (uint64_t)buf_[0] | ((uint64_t)buf_[1] << 32);
        }
        const uint64_t lo = buf_[63];
        refill();
        index_ = 1;
        return lo | ((uint64_t)buf_[0] << 32);
    }

private:
    static uint32_t rotl(uint32_t v, int k) { return (v << k) | (v >> (32 - k)); }
    static void quarter(uint32_t *x, int a, int b, int c, int d)
    {
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
    }
    void refill()
    {
        for (int blk = 0; blk < 4; ++blk) {
            const uint64_t ctr = counter_ + (uint64_t)blk;
            const uint32_t init[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                                       key_[0], key_[1], key_[2], key_[3],
                                       key_[4], key_[5], key_[6], key_[7],
                                       (uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u};
            uint32_t x[16];
            std::memcpy(x, init, sizeof x);
            for (int dr = 0; dr < 6; ++dr) {  // 12 rounds
                quarter(x, 0, 4, 8, 12); quarter(x, 1, 5, 9, 13);
                quarter(x, 2, 6, 10, 14); quarter(x, 3, 7, 11, 15);
                quarter(x, 0, 5, 10, 15); quarter(x, 1, 6, 11, 12);
                quarter(x, 2, 7, 8, 13); quarter(x, 3, 4, 9, 14);
            }
            for (int i = 0; i < 16; ++i) buf_[16 * blk + i] = x[i] + init[i];
        }
        counter_ += 4;
    }

    uint32_t key_[8];
    uint32_t buf_[64];
    uint64_t counter_ = 0;
    int index_ = 64;
};
}  // namespace

Plane::Plane(size_t w, size_t h, std::vector<float> d) : width(w), height(h), data(std::move(d))
{
    if (data.size() != w * h) throw std::runtime_error("plane size mismatch");
}

// src/model.rs:60-67
float Plane::get_clamped(long long x, long long y) const
{
    if (is_empty()) return 0.0f;
    const long long xi = std::min(std::max(x, 0LL), (long long)width - 1);
    const long long yi = std::min(std::max(y, 0LL), (long long)height - 1);
    return get((size_t)xi, (size_t)yi);
}

// src/model.rs:77-98
Plane Plane::resize_nearest(size_t nw, size_t nh) const
{
    if (nw == width && nh == height) return *this;
    if (nw == 0 || nh == 0 || width == 0 || height == 0) return Plane(nw, nh);
    Plane r(nw, nh);
    const float sx = (float)width / (float)nw, sy = (float)height / (float)nh;
    for (size_t y = 0; y < nh; ++y) {
        const float fy = std::min(std::max(((float)y + 0.5f) * sy - 0.5f, 0.0f), (float)(height - 1));
        const size_t iy = (size_t)std::round(fy);
        for (size_t x = 0; x < nw; ++x) {
            const float fx = std::min(std::max(((float)x + 0.5f) * sx - 0.5f, 0.0f), (float)(width - 1));
            r.set(x, y, get((size_t)std::round(fx), iy));
        }
    }
    return r;
}

RadiusProfile RadiusProfile::from_params(const Params &p)
{
    RadiusProfile r;
    r.dist = p.radius_dist;
    r.mean_linear = p.radius_mean;
    if (p.radius_log_mu) r.log_mu = (double)*p.radius_log_mu;
    if (p.radius_log_sigma) r.log_sigma = (double)*p.radius_log_sigma;
    if (p.radius_dist == RadiusDist::Lognorm) {
        if (!r.log_mu) throw std::runtime_error("missing log-normal mean; parameters were not derived");
        if (!r.log_sigma) throw std::runtime_error("missing log-normal sigma; parameters were not derived");
        if (!(*r.log_sigma >= 0.0)) throw std::runtime_error("invalid standard deviation");
    }
    return r;
}

// src/model.rs:150-164
float RadiusProfile::quantile(float p) const
{
    if (dist == RadiusDist::Const) return mean_linear;
    const double mu = log_mu ? *log_mu : (double)mean_linear;
    const double sigma = log_sigma ? *log_sigma : 0.0;
    if (sigma == 0.0) return (float)std::exp(mu);
    return (float)std::exp(mu + sigma * inverse_normal_cdf((double)p));
}

// Standard normal quantile: the Abramowitz-Stegun 26.2.23 starting value
// refined by Halley steps on erfc, accurate to a few ulp of double.
double inverse_normal_cdf(double p)
{
    if (!(p > 0.0)) return -INFINITY;
    if (!(p < 1.0)) return INFINITY;
    const double q = p < 0.5 ? p : 1.0 - p;
    const double t = std::sqrt(-2.0 * std::log(q));
    double x = t - (2.515517 + 0.802853 * t + 0.010328 * t * t) /
                       (1.0 + 1.432788 * t + 0.189269 * t * t + 0.001308 * t * t * t);
    if (p < 0.5) x = -x;
    for (int it = 0; it < 4; ++it) {
        // fx = Phi(x) - p, evaluated on the tail that keeps precision
        const double fx = x < 0.0 ? 0.5 * std::erfc(-x / std::sqrt(2.0)) - p
                                  : (1.0 - p) - 0.5 * std::erfc(x / std::sqrt(2.0));
        const double pdf = std::exp(-0.5 * x * x) / std::sqrt(2.0 * 3.14159265358979323846);
        const double u = fx / pdf;
        x = x - u / (1.0 + 0.5 * x * u);
    }
    return x;
}

// src/model.rs:267-298
std::pair<size_t, size_t> resolve_output_size(const Params &p, std::pair<size_t, size_t> in)
{
    if (p.size) {
        const size_t w = p.size->first;
        if (w == 0) throw std::runtime_error("output width must be > 0");
        size_t h;
        if (p.size->second) {
            h = *p.size->second;
        } else {
            const float iw = (float)in.first, ih = (float)in.second;
            if (iw <= 0.0f || ih <= 0.0f) throw std::runtime_error("could not infer output height");
            const float aspect = ih / iw;
            h = (size_t)std::max(std::round((float)w * aspect), 1.0f);
        }
        if (h == 0) throw std::runtime_error("output height must be > 0");
        return {w, h};
    }
    const float w = std::max(std::ceil((float)in.first * p.zoom), 1.0f);
    const float h = std::max(std::ceil((float)in.second * p.zoom), 1.0f);
    return {(size_t)w, (size_t)h};
}

// src/rng.rs:9-24
std::vector<float> make_offsets(uint64_t seed, size_t n, float sigma)
{
    std::vector<float> out(2 * n);
    if (n == 0) return out;
    const double s = (double)max_nan(sigma, 1.1920928955078125e-07f);
    StdRng rng(fg::mix(seed, fg::kOffsetStream));
    auto next = [&rng]() { return rng.next_u64(); };
    for (size_t k = 0; k < n; ++k) {
        const double x = 0.0 + s * fg::standard_normal(next, kZX, kZF);
        const double y = 0.0 + s * fg::standard_normal(next, kZX, kZF);
        out[2 * k] = (float)x;
        out[2 * k + 1] = (float)y;
    }
    return out;
}

// src/model.rs:181-226
Derived derive_common(const Params &p, std::pair<size_t, size_t> in)
{
    if (in.first == 0 || in.second == 0) throw std::runtime_error("input image is empty after ROI");
    const auto out = resolve_output_size(p, in);
    if (out.first == 0 || out.second == 0) throw std::runtime_error("output dimensions must be positive");
    Derived d;
    d.input_width = in.first;
    d.input_height = in.second;
    d.output_width = out.first;
    d.output_height = out.second;
    const float mean_sq = p.radius_mean * p.radius_mean;
    const float variance = p.radius_stddev * p.radius_stddev;
    d.inv_e_pi_r2 = 1.0f / (kPiF * max_nan(mean_sq + variance, kEpsilon));
    d.radius = RadiusProfile::from_params(p);
    const float rm = p.max_radius.kind == MaxRadius::Absolute ? p.max_radius.value
                                                               : d.radius.quantile(p.max_radius.value);
    d.rm = max_nan(rm, kEpsilon);
    d.delta = max_nan(p.cell_delta ? *p.cell_delta : default_cell_delta(p.radius_mean), kEpsilon);
    d.offsets = make_offsets(p.seed, p.n_samples, p.sigma_px);
    d.offsets_input.resize(d.offsets.size());
    for (size_t k = 0; k < d.offsets.size(); ++k) d.offsets_input[k] = d.offsets[k] / p.zoom;
    return d;
}

// src/model.rs:228-250
std::pair<Plane, float> normalize_plane(const Plane &plane)
{
    float mx = 0.0f;
    for (float v : plane.data) mx = max_nan(mx, v);
    const bool already = mx <= 1.0f + kEpsilon;
    Plane out(plane.width, plane.height);
    const float denom = max_nan(mx + kEpsilon, kEpsilon);
    for (size_t k = 0; k < plane.data.size(); ++k) {
        float v = already ? plane.data[k] : plane.data[k] / denom;
        if (v < 0.0f) v = 0.0f;                       // f32::clamp keeps NaN
        if (v > 1.0f - kEpsilon) v = 1.0f - kEpsilon;
        out.data[k] = v;
    }
    return {std::move(out), mx};
}

// src/model.rs:252-265
Plane lambda_plane(const Plane &normalized, float inv_e_pi_r2)
{
    Plane out(normalized.width, normalized.height);
    for (size_t k = 0; k < normalized.data.size(); ++k) {
        float c = normalized.data[k];
        if (c < 0.0f) c = 0.0f;
        if (c > 1.0f - kEpsilon) c = 1.0f - kEpsilon;
        const float safe = max_nan(1.0f - c, kEpsilon);
        const float activity = -inv_e_pi_r2 * std::log(safe);
        out.data[k] = activity < kMaxLambda ? activity : kMaxLambda;  // f32::min ignores NaN
    }
    return out;
}

// src/choose.rs:4-26
Algo choose_algorithm(const Params &p, const Derived &d)
{
    if (p.algo != Algo::Auto) return p.algo;
    const float mean = max_nan(p.radius_mean, 1e-6f);
    const float sigma_ratio = mean > 0.0f ? p.radius_stddev / mean : 0.0f;
    const float rm_ratio = d.rm / mean;
    if (sigma_ratio < 0.1f && mean < 0.5f && p.n_samples <= 64) return Algo::Pixel;
    if (rm_ratio > 8.0f || sigma_ratio > 0.6f || p.n_samples > 96) return Algo::Grain;
    if (p.n_samples <= 24 && rm_ratio < 5.0f) return Algo::Pixel;
    return Algo::Grain;
}

// src/lib.rs:205-222
RenderStats make_stats(const Params &p, const Derived &d, Algo algorithm)
{
    const float mean = max_nan(p.radius_mean, 1e-6f);
    RenderStats s;
    s.algorithm = algorithm;
    s.device = p.device;
    s.input_size = {d.input_width, d.input_height};
    s.output_size = {d.output_width, d.output_height};
    s.n_samples = p.n_samples;
    s.sigma_ratio = mean > 0.0f ? p.radius_stddev / mean : 0.0f;
    s.rm_ratio = d.rm / mean;
    return s;
}

}  // namespace film_grain
