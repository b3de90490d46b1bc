// fg_host_capi.cpp -- C ABI of the host library (include/film_grain.h).
#include <cmath>
#include <cstring>

#include "../../../include/film_grain.h"
#include "fg_host.h"

using namespace film_grain;

namespace {

void put_err(char *err, size_t len, const std::string &m)
{
    if (err && len) {
        std::strncpy(err, m.c_str(), len - 1);
        err[len - 1] = '\0';
    }
}

CliArgs to_cli(const fgh_cli_args *a)
{
    CliArgs c;
    if (a->input_path) c.input_path = a->input_path;
    if (a->output_path) c.output_path = a->output_path;
    c.radius_dist = a->radius_dist == 1 ? RadiusDist::Lognorm : RadiusDist::Const;
    c.radius_mean = a->radius_mean;
    c.radius_stddev = a->radius_stddev;
    c.zoom = a->zoom;
    c.sigma_px = a->sigma_px;
    c.n_samples = a->n_samples;
    c.algo = a->algo == 1 ? Algo::Grain : (a->algo == 2 ? Algo::Pixel : Algo::Auto);
    if (a->max_radius) c.max_radius = a->max_radius;
    if (a->cell) c.cell = a->cell;
    c.color_mode = a->color_mode == 1 ? ColorMode::Rgb : ColorMode::Luma;
    c.device = a->device == 0 ? Device::Cpu : (a->device == 1 ? Device::Gpu : Device::Hip);
    if (a->roi) c.roi = std::string(a->roi);
    if (a->size) c.size = std::string(a->size);
    c.seed = (uint32_t)a->seed;
    c.dry_run = a->dry_run != 0;
    c.explain = a->explain != 0;
    if (a->output_format) c.output_format = std::string(a->output_format);
    return c;
}

Params to_params(const fgh_cli_args *a)
{
    ParamsBuilder b = params_builder_from_cli(to_cli(a));
    b.seed = a->seed;  // keep all 64 bits (Params::seed is u64)
    return b.build();
}

void fill(fgh_derived *o, const Params &p, const Derived &d, Algo algo)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->in_w = (uint32_t)d.input_width;
    o->in_h = (uint32_t)d.input_height;
    o->out_w = (uint32_t)d.output_width;
    o->out_h = (uint32_t)d.output_height;
    o->n_samples = p.n_samples;
    o->algorithm = algo == Algo::Grain ? 1u : 2u;
    o->dist_kind = p.radius_dist == RadiusDist::Lognorm ? 1u : 0u;
    o->seed = p.seed;
    o->zoom = p.zoom;
    o->delta = d.delta;
    o->rm = d.rm;
    o->radius_mean = p.radius_mean;
    o->radius_stddev = p.radius_stddev;
    o->log_mu = p.radius_log_mu ? *p.radius_log_mu : 0.0f;
    o->log_sigma = p.radius_log_sigma ? *p.radius_log_sigma : 0.0f;
    o->inv_e_pi_r2 = d.inv_e_pi_r2;
    const RenderStats s = make_stats(p, d, algo);
    o->sigma_ratio = s.sigma_ratio;
    o->rm_ratio = s.rm_ratio;
}

int code_of(const RenderError &e)
{
    switch (e.kind) {
    case RenderError::ParamsE: return FGH_EPARAMS;
    case RenderError::Image: return FGH_EIMAGE;
    case RenderError::Io: return FGH_EIO;
    case RenderError::Gpu: return FGH_EGPU;
    case RenderError::Unsupported: return FGH_EUNSUPPORTED;
    case RenderError::Cancelled: return FGH_ECANCELLED;
    case RenderError::Message: return FGH_EMESSAGE;
    }
    return FGH_EMESSAGE;
}

template <class F>
int guarded(char *err, size_t errlen, F &&f)
{
    try {
        return f();
    } catch (const ParamsError &e) {
        put_err(err, errlen, std::string("parameter error: ") + e.what());
        return FGH_EPARAMS;
    } catch (const RenderError &e) {
        put_err(err, errlen, e.what());
        return code_of(e);
    } catch (const std::exception &e) {
        put_err(err, errlen, e.what());
        return FGH_EMESSAGE;
    }
}

}  // namespace

extern "C" {

int fgh_derive(const fgh_cli_args *args, uint32_t in_w, uint32_t in_h, fgh_derived *out, float *offsets_xy,
               float *offsets_input_xy, char *err, size_t errlen)
{
    return guarded(err, errlen, [&]() {
        const Params p = to_params(args);
        Derived d;
        try {
            d = derive_common(p, {in_w, in_h});
        } catch (const std::exception &e) {
            throw RenderError(RenderError::Message, e.what());
        }
        fill(out, p, d, choose_algorithm(p, d));
        if (offsets_xy) std::memcpy(offsets_xy, d.offsets.data(), d.offsets.size() * sizeof(float));
        if (offsets_input_xy)
            std::memcpy(offsets_input_xy, d.offsets_input.data(), d.offsets_input.size() * sizeof(float));
        return FGH_OK;
    });
}

int fgh_make_offsets(uint64_t seed, uint32_t n, float sigma, float *out_xy)
{
    const std::vector<float> o = make_offsets(seed, n, sigma);
    if (n) std::memcpy(out_xy, o.data(), o.size() * sizeof(float));
    return FGH_OK;
}

float fgh_normalize_plane(const float *src, size_t n, float *dst)
{
    Plane p(n, 1, std::vector<float>(src, src + n));
    auto r = normalize_plane(p);
    std::memcpy(dst, r.first.data.data(), n * sizeof(float));
    return r.second;
}

void fgh_lambda_plane(const float *normalized, size_t n, float inv_e_pi_r2, float *dst)
{
    Plane p(n, 1, std::vector<float>(normalized, normalized + n));
    const Plane l = lambda_plane(p, inv_e_pi_r2);
    std::memcpy(dst, l.data.data(), n * sizeof(float));
}

double fgh_inverse_normal_cdf(double p) { return inverse_normal_cdf(p); }

int fgh_render_rgb(const fgh_cli_args *args, const float *rgb, uint32_t w, uint32_t h, uint8_t *out_rgb8,
                   size_t out_rgb8_cap, float *out_planes, size_t out_planes_cap, fgh_derived *stats,
                   fgh_cancel_fn cancel, void *cancel_user, char *err, size_t errlen)
{
    return guarded(err, errlen, [&]() {
        const Params p = to_params(args);
        const InputImage in = InputImage::from_rgb_f32(rgb, w, h, p.color_mode, p.roi);
        CancelCheck cc;
        if (cancel) cc = [cancel, cancel_user]() { return cancel(cancel_user) != 0; };
        std::vector<Plane> planes;
        auto r = render_workspace(in.to_workspace(), p, cancel ? &cc : nullptr, &planes);
        const Derived d = derive_common(p, in.dimensions());
        fill(stats, p, d, r.second.algorithm);
        if (out_rgb8) {
            if (out_rgb8_cap < r.first.data.size()) {
                put_err(err, errlen, "output buffer too small");
                return FGH_ECAPACITY;
            }
            std::memcpy(out_rgb8, r.first.data.data(), r.first.data.size());
        }
        if (out_planes) {
            size_t need = 0;
            for (const auto &pl : planes) need += pl.data.size();
            if (out_planes_cap < need) {
                put_err(err, errlen, "plane buffer too small");
                return FGH_ECAPACITY;
            }
            size_t off = 0;
            for (const auto &pl : planes) {
                std::memcpy(out_planes + off, pl.data.data(), pl.data.size() * sizeof(float));
                off += pl.data.size();
            }
        }
        return FGH_OK;
    });
}

int fgh_render_file(const fgh_cli_args *args, fgh_derived *stats, char *err, size_t errlen)
{
    return guarded(err, errlen, [&]() {
        const Params p = to_params(args);
        const InputImage in = InputImage::from_path(p.input_path, p.color_mode, p.roi);
        auto r = render_workspace(in.to_workspace(), p, nullptr, nullptr);
        save_image(r.first, p.output_path, resolve_format(p));
        fill(stats, p, derive_common(p, in.dimensions()), r.second.algorithm);
        return FGH_OK;
    });
}

int fgh_render_rgb_frame(const fgh_cli_args *args, const float *rgb, uint32_t w, uint32_t h, uint8_t *out_rgb8,
                         size_t out_rgb8_cap, float *out_planes, size_t out_planes_cap, fgh_derived *stats,
                         fgh_cancel_fn cancel, void *cancel_user, char *err, size_t errlen)
{
    return guarded(err, errlen, [&]() {
        const Params p = to_params(args);
        CancelCheck cc;
        if (cancel) cc = [cancel, cancel_user]() { return cancel(cancel_user) != 0; };
        std::vector<Plane> planes;
        auto r = render_rgb_frame(rgb, w, h, p, cancel ? &cc : nullptr, out_planes ? &planes : nullptr);
        const std::pair<size_t, size_t> dims =
            p.roi ? std::pair<size_t, size_t>(p.roi->x1 - p.roi->x0, p.roi->y1 - p.roi->y0)
                  : std::pair<size_t, size_t>(w, h);
        fill(stats, p, derive_common(p, dims), r.second.algorithm);
        if (out_rgb8) {
            if (out_rgb8_cap < r.first.data.size()) {
                put_err(err, errlen, "output buffer too small");
                return FGH_ECAPACITY;
            }
            std::memcpy(out_rgb8, r.first.data.data(), r.first.data.size());
        }
        if (out_planes) {
            size_t need = 0;
            for (const auto &pl : planes) need += pl.data.size();
            if (out_planes_cap < need) {
                put_err(err, errlen, "plane buffer too small");
                return FGH_ECAPACITY;
            }
            size_t off = 0;
            for (const auto &pl : planes) {
                std::memcpy(out_planes + off, pl.data.data(), pl.data.size() * sizeof(float));
                off += pl.data.size();
            }
        }
        return FGH_OK;
    });
}

void fgh_logf_array(const float *x, size_t n, float *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = std::log(x[i]);
}

int fgh_color_split(const float *rgb, uint32_t w, uint32_t h, uint32_t color_mode, float *p0, float *p1, float *p2)
{
    return guarded(nullptr, 0, [&]() {
        const Workspace ws = Workspace::from_rgb_f32(rgb, w, h, color_mode == 1 ? ColorMode::Rgb : ColorMode::Luma);
        const size_t n = (size_t)w * h;
        const Plane *src[3] = {ws.kind == Workspace::Luma ? &ws.y : &ws.planes[0],
                               ws.kind == Workspace::Luma ? &ws.cb : &ws.planes[1],
                               ws.kind == Workspace::Luma ? &ws.cr : &ws.planes[2]};
        float *dst[3] = {p0, p1, p2};
        for (int i = 0; i < 3; ++i) std::memcpy(dst[i], src[i]->data.data(), n * sizeof(float));
        return FGH_OK;
    });
}

int fgh_color_roundtrip(const float *rgb, uint32_t w, uint32_t h, uint32_t color_mode, uint8_t *out_rgb8)
{
    return guarded(nullptr, 0, [&]() {
        const Workspace ws = Workspace::from_rgb_f32(rgb, w, h, color_mode == 1 ? ColorMode::Rgb : ColorMode::Luma);
        const RgbImage img = ws.into_rgb_image();
        std::memcpy(out_rgb8, img.data.data(), img.data.size());
        return FGH_OK;
    });
}

int fgh_load_image(const char *path, uint32_t *w, uint32_t *h, float *out_rgb, size_t cap, char *err,
                   size_t errlen)
{
    return guarded(err, errlen, [&]() {
        size_t iw = 0, ih = 0;
        const std::vector<float> rgb = load_image_rgb_f32(path, &iw, &ih);
        *w = (uint32_t)iw;
        *h = (uint32_t)ih;
        if (out_rgb) {
            if (cap < rgb.size()) {
                put_err(err, errlen, "image buffer too small");
                return FGH_ECAPACITY;
            }
            std::memcpy(out_rgb, rgb.data(), rgb.size() * sizeof(float));
        }
        return FGH_OK;
    });
}

int fgh_save_rgb8(const char *path, const uint8_t *rgb8, uint32_t w, uint32_t h, const char *format, char *err,
                  size_t errlen)
{
    return guarded(err, errlen, [&]() {
        RgbImage img;
        img.width = w;
        img.height = h;
        img.data.assign(rgb8, rgb8 + (size_t)w * h * 3);
        save_image(img, path, format ? format : "png");
        return FGH_OK;
    });
}

int fgh_dry_run(const fgh_cli_args *args, fgh_derived *stats, char *err, size_t errlen)
{
    return guarded(err, errlen, [&]() {
        const Params p = to_params(args);
        const InputImage in = InputImage::from_path(p.input_path, p.color_mode, p.roi);
        const RenderStats s = dry_run_with_input_image(in, p);
        fill(stats, p, derive_common(p, in.dimensions()), s.algorithm);
        return FGH_OK;
    });
}

}  // extern "C"
