// fg_host.h -- C++ host library of the film-grain backend.
//
// Restates the reference library's host surface so a caller of the Rust crate
// finds the same entry points, types and routing:
//   params   RadiusDist / Algo / ColorMode / Device, Roi, MaxRadius, Params,
//            ParamsBuilder::build, CliArgs -> ParamsBuilder, build_params,
//            default_cell_delta, the string parsers   (src/params.rs)
//   model    Plane, RadiusProfile, Derived, derive_common, normalize_plane,
//            lambda_plane, resolve_output_size          (src/model.rs)
//   rng      make_offsets (StdRng = ChaCha12 + Normal)  (src/rng.rs:9-24)
//   routing  choose_algorithm                           (src/choose.rs)
//   colour   Workspace, InputImage, into_rgb_image      (src/color.rs)
//   library  render, render_to_image, render_with_input_image[_cancelable],
//            dry_run*, RenderStats, RenderError         (src/lib.rs)
// Device::Hip (and Device::Gpu, its alias) renders through the C ABI of
// include/film_grain_hip.h.  The CPU integrators are not part of this backend:
// Device::Cpu returns RenderError::Unsupported.
#ifndef FG_HOST_H
#define FG_HOST_H

#include <cstdint>
#include <functional>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace film_grain {

enum class RadiusDist { Const, Lognorm };
enum class Algo { Auto, Grain, Pixel };
enum class ColorMode { Luma, Rgb };
enum class Device { Cpu, Gpu, Hip };

struct Roi {
    uint32_t x0, y0, x1, y1;
};

struct MaxRadius {
    enum Kind { Absolute, Quantile } kind = Quantile;
    float value = 0.999f;
};

// src/params.rs:116-139
struct ParamsError : std::runtime_error {
    std::string field;
    std::string message;
    ParamsError(std::string f, std::string m)
        : std::runtime_error(f + ": " + m), field(std::move(f)), message(std::move(m)) {}
};

// src/params.rs:45-68
struct Params {
    std::string input_path, output_path;
    RadiusDist radius_dist = RadiusDist::Const;
    float radius_mean = 0.1f;
    float radius_stddev = 0.0f;
    std::optional<float> radius_log_mu, radius_log_sigma;
    float zoom = 1.0f;
    float sigma_px = 0.8f;
    uint32_t n_samples = 32;
    Algo algo = Algo::Auto;
    MaxRadius max_radius;
    std::optional<float> cell_delta;
    ColorMode color_mode = ColorMode::Luma;
    std::optional<Roi> roi;
    std::optional<std::pair<uint32_t, std::optional<uint32_t>>> size;
    uint64_t seed = 5489;
    bool dry_run = false, explain = false;
    Device device = Device::Cpu;
    std::optional<std::string> output_format;
};

// src/params.rs:70-91, 141-180
struct ParamsBuilder {
    std::string input_path, output_path;
    RadiusDist radius_dist = RadiusDist::Const;
    float radius_mean = 0.1f;
    float radius_stddev = 0.0f;
    float zoom = 1.0f;
    float sigma_px = 0.8f;
    uint32_t n_samples = 32;
    Algo algo = Algo::Auto;
    MaxRadius max_radius;
    std::optional<float> cell_delta;
    ColorMode color_mode = ColorMode::Luma;
    std::optional<Roi> roi;
    std::optional<std::pair<uint32_t, std::optional<uint32_t>>> size;
    uint64_t seed = 5489;
    bool dry_run = false, explain = false;
    Device device = Device::Cpu;
    std::optional<std::string> output_format;

    Params build() const;  // throws ParamsError
};

// src/params.rs:93-114 (CLI-level values: f64 numbers and strings)
struct CliArgs {
    std::string input_path, output_path;
    RadiusDist radius_dist = RadiusDist::Const;
    double radius_mean = 0.10, radius_stddev = 0.0, zoom = 1.0, sigma_px = 0.8;
    uint32_t n_samples = 32;
    Algo algo = Algo::Auto;
    std::string max_radius = "quantile 0.999";
    std::string cell = "approx mu_r";
    ColorMode color_mode = ColorMode::Luma;
    Device device = Device::Cpu;
    std::optional<std::string> roi, size;
    uint32_t seed = 5489;
    bool dry_run = false, explain = false;
    std::optional<std::string> output_format;
};

ParamsBuilder params_builder_from_cli(const CliArgs &args);  // TryFrom<CliArgs>
Params build_params(const CliArgs &args);
float default_cell_delta(float radius_mean);
// derive_radius_parameters (src/params.rs:223-253): (stddev, log_mu, log_sigma)
void derive_radius_parameters(float mean, float stddev, RadiusDist dist, float *out_stddev,
                              std::optional<float> *log_mu, std::optional<float> *log_sigma);
MaxRadius parse_max_radius(const std::string &value);
std::optional<float> parse_cell_delta(const std::string &cell, float radius_mean);
std::optional<Roi> parse_roi(const std::optional<std::string> &raw);
std::optional<std::pair<uint32_t, std::optional<uint32_t>>> parse_size(const std::optional<std::string> &raw);

// src/model.rs:13-99
struct Plane {
    size_t width = 0, height = 0;
    std::vector<float> data;
    Plane() = default;
    Plane(size_t w, size_t h) : width(w), height(h), data(w * h, 0.0f) {}
    Plane(size_t w, size_t h, std::vector<float> d);
    bool is_empty() const { return data.empty(); }
    float get(size_t x, size_t y) const { return data[y * width + x]; }
    void set(size_t x, size_t y, float v) { data[y * width + x] = v; }
    float get_clamped(long long x, long long y) const;
    Plane resize_nearest(size_t new_width, size_t new_height) const;
};

// src/model.rs:101-165
struct RadiusProfile {
    RadiusDist dist = RadiusDist::Const;
    float mean_linear = 0.0f;
    std::optional<double> log_mu, log_sigma;
    static RadiusProfile from_params(const Params &p);
    float quantile(float p) const;
};

// src/model.rs:167-179
struct Derived {
    size_t input_width = 0, input_height = 0, output_width = 0, output_height = 0;
    float inv_e_pi_r2 = 0.0f, rm = 0.0f, delta = 0.0f;
    std::vector<float> offsets;        // N (x, y) pairs, output pixels
    std::vector<float> offsets_input;  // offsets / zoom, input pixels
    RadiusProfile radius;
};

Derived derive_common(const Params &p, std::pair<size_t, size_t> input_size);  // throws std::runtime_error
std::pair<Plane, float> normalize_plane(const Plane &plane);
Plane lambda_plane(const Plane &normalized, float inv_e_pi_r2);
std::pair<size_t, size_t> resolve_output_size(const Params &p, std::pair<size_t, size_t> input_size);
std::vector<float> make_offsets(uint64_t seed, size_t n, float sigma);
Algo choose_algorithm(const Params &p, const Derived &d);
double inverse_normal_cdf(double p);

// src/lib.rs:28-55
struct RenderError : std::runtime_error {
    enum Kind { ParamsE, Image, Io, Gpu, Unsupported, Cancelled, Message } kind;
    RenderError(Kind k, const std::string &m) : std::runtime_error(m), kind(k) {}
};

struct RenderStats {
    Algo algorithm;
    Device device;
    std::pair<size_t, size_t> input_size, output_size;
    uint32_t n_samples;
    float sigma_ratio, rm_ratio;
};

RenderStats make_stats(const Params &p, const Derived &d, Algo algorithm);

// src/color.rs
struct RgbImage {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> data;  // interleaved RGB8
};

class Workspace {
public:
    enum Kind { Luma, Rgb } kind = Luma;
    Plane y, cb, cr;   // Luma
    Plane planes[3];   // Rgb
    // interleaved RGB f32 in [0,1] (as image::to_rgb32f), cropped already
    static Workspace from_rgb_f32(const float *rgb, size_t w, size_t h, ColorMode mode);
    std::pair<size_t, size_t> dimensions() const;
    void for_each_plane(const std::function<Plane(const Plane &, size_t)> &f);
    RgbImage into_rgb_image() const;
};

class InputImage {
public:
    ColorMode color_mode = ColorMode::Luma;
    Workspace ws;
    static InputImage from_path(const std::string &path, ColorMode mode, const std::optional<Roi> &roi);
    static InputImage from_rgb_f32(const float *rgb, size_t w, size_t h, ColorMode mode,
                                   const std::optional<Roi> &roi);
    Workspace to_workspace() const { return ws; }
    std::pair<size_t, size_t> dimensions() const { return ws.dimensions(); }
};

using CancelCheck = std::function<bool()>;

// src/lib.rs:57-112
RenderStats render(const Params &p);
std::pair<RgbImage, RenderStats> render_to_image(const Params &p);
std::pair<RgbImage, RenderStats> render_with_input_image(const InputImage &input, const Params &p);
std::pair<RgbImage, RenderStats> render_with_input_image_cancelable(const InputImage &input,
                                                                    const Params &p,
                                                                    const CancelCheck &cancel);
RenderStats dry_run(const Params &p);
RenderStats dry_run_with_input_image(const InputImage &input, const Params &p);
RenderStats dry_run_with_input_image_cancelable(const InputImage &input, const Params &p,
                                                const CancelCheck &cancel);
// Renders a workspace and also hands back the rendered f32 planes.
std::pair<RgbImage, RenderStats> render_workspace(Workspace ws, const Params &p,
                                                  const CancelCheck *cancel,
                                                  std::vector<Plane> *planes_out);
// The same result from an interleaved RGB f32 image with the whole frame on
// the device (colour split, lambda pre-pass, integrator, merge: one upload,
// one download, fg_render_frame); the ROI crop stays on the host.
std::pair<RgbImage, RenderStats> render_rgb_frame(const float *rgb, size_t w, size_t h, const Params &p,
                                                  const CancelCheck *cancel, std::vector<Plane> *planes_out);

// image I/O (PNG 8-bit grey/RGB/RGBA via zlib, binary PGM/PPM, PFM)
std::vector<float> load_image_rgb_f32(const std::string &path, size_t *w, size_t *h);
void save_image(const RgbImage &img, const std::string &path, const std::string &format);
std::string resolve_format(const Params &p);

const char *algo_name(Algo a);
const char *device_name(Device d);

}  // namespace film_grain

#endif
