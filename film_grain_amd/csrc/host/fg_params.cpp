// fg_params.cpp -- parameter types, validation and parsing (src/params.rs).
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <limits>

#include "fg_host.h"

namespace film_grain {

namespace {

std::string trim(const std::string &s)
{
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

std::string lower(std::string s)
{
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// str::parse::<f32>: the whole string must be a float literal (inf/nan accepted).
bool parse_f32(const std::string &s, float *out)
{
    if (s.empty() || std::isspace((unsigned char)s[0])) return false;
    const char *b = s.c_str();
    char *end = nullptr;
    errno = 0;
    const float v = std::strtof(b, &end);
    if (end != b + s.size()) return false;
    // Rust does not accept hexadecimal floats
    if (s.find_first_of("xX") != std::string::npos && lower(s).find("inf") == std::string::npos)
        return false;
    *out = v;
    return true;
}

// str::parse::<u32>: optional '+', decimal digits only.
bool parse_u32_strict(const std::string &s, uint32_t *out)
{
    size_t i = 0;
    if (i < s.size() && s[i] == '+') ++i;
    if (i == s.size()) return false;
    uint64_t v = 0;
    for (; i < s.size(); ++i) {
        if (!std::isdigit((unsigned char)s[i])) return false;
        v = v * 10 + (uint64_t)(s[i] - '0');
        if (v > 0xFFFFFFFFull) return false;
    }
    *out = (uint32_t)v;
    return true;
}

uint32_t parse_u32(const std::string &raw, const char *field)
{
    uint32_t v;
    if (!parse_u32_strict(trim(raw), &v))
        throw ParamsError(field, "expected an unsigned integer without suffixes");
    return v;
}

float ensure_positive(float value, const char *field)
{
    if (!std::isfinite(value)) throw ParamsError(field, "value must be finite");
    if (value <= 0.0f) throw ParamsError(field, "value must be greater than 0");
    return value;
}

float ensure_non_negative(float value, const char *field)
{
    if (!std::isfinite(value)) throw ParamsError(field, "value must be finite");
    if (value < 0.0f) throw ParamsError(field, "value must be >= 0");
    return value;
}

std::optional<float> ensure_cell_delta(std::optional<float> value, float radius_mean)
{
    if (value) {
        if (!std::isfinite(*value)) throw ParamsError("cell", "cell size must be finite");
        if (*value <= 0.0f) throw ParamsError("cell", "cell size must be > 0");
        return value;
    }
    return default_cell_delta(radius_mean);
}

MaxRadius ensure_max_radius(MaxRadius spec)
{
    if (spec.kind == MaxRadius::Absolute) {
        if (!std::isfinite(spec.value)) throw ParamsError("max-radius", "absolute radius must be finite");
        if (spec.value <= 0.0f) throw ParamsError("max-radius", "absolute radius must be > 0");
    } else {
        if (!std::isfinite(spec.value)) throw ParamsError("max-radius", "quantile must be finite");
        if (!(0.0f < spec.value && spec.value < 1.0f))
            throw ParamsError("max-radius", "quantile must lie in the open interval (0,1)");
    }
    return spec;
}

std::optional<Roi> ensure_roi(const std::optional<Roi> &roi)
{
    if (roi && (roi->x1 <= roi->x0 || roi->y1 <= roi->y0))
        throw ParamsError("roi", "roi end must be greater than start (exclusive bounds)");
    return roi;
}

std::optional<std::pair<uint32_t, std::optional<uint32_t>>> ensure_size(
    const std::optional<std::pair<uint32_t, std::optional<uint32_t>>> &size)
{
    if (size) {
        if (size->first == 0) throw ParamsError("size", "output width must be > 0");
        if (size->second && *size->second == 0) throw ParamsError("size", "output height must be > 0");
    }
    return size;
}

float to_positive_f32(double value, const char *field)
{
    if (!std::isfinite(value)) throw ParamsError(field, "value must be finite");
    if (value <= 0.0) throw ParamsError(field, "value must be greater than 0");
    if (value > (double)std::numeric_limits<float>::max())
        throw ParamsError(field, "value is too large for single precision");
    return (float)value;
}

float to_non_negative_f32(double value, const char *field)
{
    if (!std::isfinite(value)) throw ParamsError(field, "value must be finite");
    if (value < 0.0) throw ParamsError(field, "value must be >= 0");
    if (value > (double)std::numeric_limits<float>::max())
        throw ParamsError(field, "value is too large for single precision");
    return (float)value;
}

std::pair<uint32_t, uint32_t> parse_pair(const std::string &raw, const char *field)
{
    const size_t c = raw.find(',');
    if (c == std::string::npos) throw ParamsError(field, "expected two comma-separated integers");
    const std::string a = raw.substr(0, c), rest = raw.substr(c + 1);
    if (rest.find(',') != std::string::npos) throw ParamsError(field, "too many commas");
    return {parse_u32(a, field), parse_u32(rest, field)};
}

}  // namespace

// src/params.rs:255-261
float default_cell_delta(float radius_mean)
{
    if (radius_mean <= 0.0f) return 1.0f;
    float inv = std::ceil(1.0f / radius_mean);
    if (!(inv >= 1.0f)) inv = 1.0f;
    return 1.0f / inv;
}

// src/params.rs:223-253; the log-normal mu and sigma are computed in f32.
void derive_radius_parameters(float mean, float stddev, RadiusDist dist, float *out_stddev,
                              std::optional<float> *log_mu, std::optional<float> *log_sigma)
{
    if (dist == RadiusDist::Const) {
        *out_stddev = 0.0f;
        log_mu->reset();
        log_sigma->reset();
        return;
    }
    if (mean <= 0.0f)
        throw ParamsError("radius", "mean radius must be positive for lognormal distribution");
    if (stddev < 0.0f) throw ParamsError("radius-stddev", "stddev must be non-negative");
    if (stddev == 0.0f) {
        *out_stddev = 0.0f;
        *log_mu = std::log(mean);
        *log_sigma = 0.0f;
        return;
    }
    const float variance_ratio = (stddev * stddev) / (mean * mean);
    const float sigma_sq = std::log(1.0f + variance_ratio);
    const float sigma = std::sqrt(sigma_sq);
    const float mu = std::log(mean) - 0.5f * sigma_sq;
    *out_stddev = stddev;
    *log_mu = mu;
    *log_sigma = sigma;
}

// src/params.rs:141-180
Params ParamsBuilder::build() const
{
    Params p;
    p.radius_mean = ensure_positive(radius_mean, "radius");
    const float sd_in = ensure_non_negative(radius_stddev, "radius-stddev");
    p.zoom = ensure_positive(zoom, "zoom");
    p.sigma_px = ensure_positive(sigma_px, "sigma");
    p.n_samples = n_samples > 1 ? n_samples : 1;
    p.max_radius = ensure_max_radius(max_radius);
    p.roi = ensure_roi(roi);
    p.size = ensure_size(size);
    p.cell_delta = ensure_cell_delta(cell_delta, p.radius_mean);
    derive_radius_parameters(p.radius_mean, sd_in, radius_dist, &p.radius_stddev, &p.radius_log_mu,
                             &p.radius_log_sigma);
    p.input_path = input_path;
    p.output_path = output_path;
    p.radius_dist = radius_dist;
    p.algo = algo;
    p.color_mode = color_mode;
    p.seed = seed;
    p.dry_run = dry_run;
    p.explain = explain;
    p.device = device;
    p.output_format = output_format;
    return p;
}

// src/params.rs:414-447
MaxRadius parse_max_radius(const std::string &value)
{
    const std::string t = trim(value);
    MaxRadius m;
    if (t.empty()) return m;
    if (lower(t).rfind("quantile", 0) == 0) {
        std::string tail = trim(t.substr(8));
        size_t k = 0;
        while (k < tail.size() && (tail[k] == '=' || tail[k] == ':')) ++k;
        tail = trim(tail.substr(k));
        if (tail.empty()) return m;
        float q;
        if (!parse_f32(tail, &q)) throw ParamsError("max-radius", "expected numeric quantile value in (0,1)");
        if (!(0.0f < q && q < 1.0f))
            throw ParamsError("max-radius", "quantile must lie in the open interval (0,1)");
        m.value = q;
        return m;
    }
    float a;
    if (!parse_f32(t, &a)) throw ParamsError("max-radius", "expected a positive number or 'quantile'");
    if (a <= 0.0f) throw ParamsError("max-radius", "absolute radius must be > 0");
    m.kind = MaxRadius::Absolute;
    m.value = a;
    return m;
}

// src/params.rs:392-412
std::optional<float> parse_cell_delta(const std::string &cell, float radius_mean)
{
    const std::string t = trim(cell);
    if (t.empty()) return std::nullopt;
    const std::string l = lower(t);
    if (l == "approx mu_r" || l == "auto" || l == "mu_r") return default_cell_delta(radius_mean);
    float v;
    if (!parse_f32(t, &v)) throw ParamsError("cell", "expected a positive number or 'approx mu_r'");
    if (v <= 0.0f) throw ParamsError("cell", "cell size must be > 0");
    return v;
}

// src/params.rs:449-472
std::optional<Roi> parse_roi(const std::optional<std::string> &raw)
{
    if (!raw) return std::nullopt;
    const std::string &text = *raw;
    const size_t c = text.find(':');
    if (c == std::string::npos) throw ParamsError("roi", "expected X0,Y0:X1,Y1");
    const std::string start = text.substr(0, c), end = text.substr(c + 1);
    if (end.find(':') != std::string::npos) throw ParamsError("roi", "unexpected extra ':'");
    const auto a = parse_pair(start, "roi");
    const auto b = parse_pair(end, "roi");
    if (b.first <= a.first || b.second <= a.second)
        throw ParamsError("roi", "roi end must be greater than start (exclusive bounds)");
    return Roi{a.first, a.second, b.first, b.second};
}

// src/params.rs:496-532
std::optional<std::pair<uint32_t, std::optional<uint32_t>>> parse_size(const std::optional<std::string> &raw)
{
    if (!raw) return std::nullopt;
    const std::string t = trim(*raw);
    if (t.empty()) throw ParamsError("size", "size string cannot be empty");
    const size_t idx = lower(t).find('x');
    if (idx != std::string::npos) {
        const std::string w = trim(t.substr(0, idx)), h = trim(t.substr(idx + 1));
        if (!w.empty() && !h.empty())
            return std::make_pair(parse_u32(w, "size"), std::optional<uint32_t>(parse_u32(h, "size")));
        if (!w.empty()) return std::make_pair(parse_u32(w, "size"), std::optional<uint32_t>());
        if (!h.empty())
            throw ParamsError("size", "height-only form is not supported; provide width or WxH");
        throw ParamsError("size", "missing width and height around separator");
    }
    return std::make_pair(parse_u32(t, "size"), std::optional<uint32_t>());
}

// src/params.rs:182-217
ParamsBuilder params_builder_from_cli(const CliArgs &a)
{
    ParamsBuilder b;
    b.radius_mean = to_positive_f32(a.radius_mean, "radius");
    b.radius_stddev = to_non_negative_f32(a.radius_stddev, "radius-stddev");
    b.zoom = to_positive_f32(a.zoom, "zoom");
    b.sigma_px = to_positive_f32(a.sigma_px, "sigma");
    b.cell_delta = parse_cell_delta(a.cell, b.radius_mean);
    b.max_radius = parse_max_radius(a.max_radius);
    b.roi = parse_roi(a.roi);
    b.size = parse_size(a.size);
    b.input_path = a.input_path;
    b.output_path = a.output_path;
    b.radius_dist = a.radius_dist;
    b.n_samples = a.n_samples;
    b.algo = a.algo;
    b.color_mode = a.color_mode;
    b.seed = (uint64_t)a.seed;
    b.dry_run = a.dry_run;
    b.explain = a.explain;
    b.device = a.device;
    b.output_format = a.output_format;
    return b;
}

Params build_params(const CliArgs &args) { return params_builder_from_cli(args).build(); }

const char *algo_name(Algo a)
{
    switch (a) {
    case Algo::Auto: return "Auto";
    case Algo::Grain: return "Grain";
    case Algo::Pixel: return "Pixel";
    }
    return "?";
}

const char *device_name(Device d)
{
    switch (d) {
    case Device::Cpu: return "Cpu";
    case Device::Gpu: return "Gpu";
    case Device::Hip: return "Hip";
    }
    return "?";
}

}  // namespace film_grain
