// fg_lib.cpp -- library entry points (src/lib.rs:57-186): derive -> choose ->
// per-plane normalise / lambda / integrate -> recombine, with the integrator
// arm running on the HIP backend through include/film_grain_hip.h.
#include <algorithm>
#include <memory>
#include <mutex>

#include "../../../include/film_grain_hip.h"
#include "fg_host.h"

namespace film_grain {

namespace {

// Process-wide HIP context, created on first use and dropped after a fatal
// error so the next render re-initialises (src/wgpu/mod.rs:51-92, 727-736).
std::mutex g_ctx_mu;
fg_ctx *g_ctx = nullptr;

fg_ctx *hip_context()
{
    std::lock_guard<std::mutex> lock(g_ctx_mu);
    if (g_ctx) return g_ctx;
    fg_ctx *c = nullptr;
    const int rc = fg_ctx_create(0, &c);
    if (rc == FG_ENODEV) throw RenderError(RenderError::Gpu, "gpu error: no compatible GPU adapter found");
    if (rc != FG_OK) throw RenderError(RenderError::Gpu, "gpu error: HIP initialisation failed");
    g_ctx = c;
    return g_ctx;
}

void invalidate_context(fg_ctx *c)
{
    std::lock_guard<std::mutex> lock(g_ctx_mu);
    if (g_ctx == c) {
        fg_ctx_destroy(g_ctx);
        g_ctx = nullptr;
    }
}

void check_cancel(const CancelCheck *cancel)
{
    if (cancel && *cancel && (*cancel)()) throw RenderError(RenderError::Cancelled, "cancelled");
}

fg_plane_desc plane_desc(const Params &p, const Derived &d)
{
    fg_plane_desc desc{};
    desc.in_w = (uint32_t)d.input_width;
    desc.in_h = (uint32_t)d.input_height;
    desc.out_w = (uint32_t)d.output_width;
    desc.out_h = (uint32_t)d.output_height;
    desc.n_samples = p.n_samples;
    desc.dist_kind = p.radius_dist == RadiusDist::Lognorm ? 1u : 0u;
    desc.seed = p.seed;
    desc.zoom = p.zoom;
    desc.delta = d.delta;
    desc.rm = d.rm;
    desc.radius_mean = p.radius_mean;
    desc.log_mu = p.radius_log_mu ? *p.radius_log_mu : 0.0f;
    desc.log_sigma = p.radius_log_sigma ? *p.radius_log_sigma : 0.0f;
    return desc;
}

struct CancelThunk {
    const CancelCheck *cancel;
    static int call(void *u)
    {
        const CancelThunk *t = static_cast<const CancelThunk *>(u);
        return (t->cancel && *t->cancel && (*t->cancel)()) ? 1 : 0;
    }
};

Plane render_plane_hip(const Plane &lambda, const Params &p, const Derived &d, Algo algo,
                       const CancelCheck *cancel)
{
    fg_ctx *ctx = hip_context();
    const fg_plane_desc desc = plane_desc(p, d);
    Plane out(d.output_width, d.output_height);
    CancelThunk thunk{cancel};
    const std::vector<float> &offs = algo == Algo::Pixel ? d.offsets_input : d.offsets;
    const int rc = algo == Algo::Pixel
                       ? fg_render_pixelwise(ctx, &desc, lambda.data.data(), offs.data(), out.data.data(),
                                             &CancelThunk::call, &thunk)
                       : fg_render_grainwise(ctx, &desc, lambda.data.data(), offs.data(), out.data.data(),
                                             &CancelThunk::call, &thunk);
    if (rc == FG_OK) return out;
    if (rc == FG_ECANCELLED) throw RenderError(RenderError::Cancelled, "cancelled");
    const std::string msg = std::string("gpu error: ") + fg_last_error(ctx);
    if (rc == FG_ENOMEM || rc == FG_EHIP) invalidate_context(ctx);
    throw RenderError(RenderError::Gpu, msg);
}

std::pair<Derived, Algo> derive_for_workspace(const Workspace &ws, const Params &p, const CancelCheck *cancel)
{
    check_cancel(cancel);
    Derived d;
    try {
        d = derive_common(p, ws.dimensions());
    } catch (const RenderError &) {
        throw;
    } catch (const std::exception &e) {
        throw RenderError(RenderError::Message, e.what());
    }
    check_cancel(cancel);
    return {std::move(d), choose_algorithm(p, d)};
}

}  // namespace

std::pair<RgbImage, RenderStats> render_workspace(Workspace ws, const Params &p, const CancelCheck *cancel,
                                                  std::vector<Plane> *planes_out)
{
    check_cancel(cancel);
    auto dv = derive_for_workspace(ws, p, cancel);
    const Derived &d = dv.first;
    const Algo algo = dv.second;
    if (p.device == Device::Cpu)
        throw RenderError(RenderError::Unsupported,
                          "unsupported: the CPU integrators are not part of this backend; use --device hip");
    hip_context();  // like wgpu::context(): fail before any plane work when there is no device
    ws.for_each_plane([&](const Plane &plane, size_t) {
        check_cancel(cancel);
        const Plane normalized = normalize_plane(plane).first;
        const Plane lambda = lambda_plane(normalized, d.inv_e_pi_r2);
        check_cancel(cancel);
        Plane out = render_plane_hip(lambda, p, d, algo, cancel);
        if (planes_out) planes_out->push_back(out);
        return out;
    });
    check_cancel(cancel);
    const RenderStats stats = make_stats(p, d, algo);
    check_cancel(cancel);
    return {ws.into_rgb_image(), stats};
}

std::pair<RgbImage, RenderStats> render_rgb_frame(const float *rgb, size_t w, size_t h, const Params &p,
                                                  const CancelCheck *cancel, std::vector<Plane> *planes_out)
{
    check_cancel(cancel);
    std::vector<float> crop;  // src/color.rs:215-231
    if (p.roi) {
        const Roi &r = *p.roi;
        if (r.x1 > w || r.y1 > h) throw RenderError(RenderError::Message, "ROI exceeds image bounds");
        const size_t cw = r.x1 - r.x0, ch = r.y1 - r.y0;
        if (cw == 0 || ch == 0) throw RenderError(RenderError::Message, "ROI width and height must be positive");
        crop.resize(cw * ch * 3);
        for (size_t yy = 0; yy < ch; ++yy)
            std::copy(rgb + ((r.y0 + yy) * w + r.x0) * 3, rgb + ((r.y0 + yy) * w + r.x0 + cw) * 3,
                      crop.begin() + yy * cw * 3);
        rgb = crop.data();
        w = cw;
        h = ch;
    }
    Derived d;
    try {
        d = derive_common(p, {w, h});
    } catch (const RenderError &) {
        throw;
    } catch (const std::exception &e) {
        throw RenderError(RenderError::Message, e.what());
    }
    const Algo algo = choose_algorithm(p, d);
    if (p.device == Device::Cpu)
        throw RenderError(RenderError::Unsupported,
                          "unsupported: the CPU integrators are not part of this backend; use --device hip");
    fg_ctx *ctx = hip_context();
    fg_frame_desc fd{};
    fd.plane = plane_desc(p, d);
    fd.color_mode = p.color_mode == ColorMode::Rgb ? FG_COLOR_RGB : FG_COLOR_LUMA;
    fd.algo = algo == Algo::Pixel ? FG_ALGO_PIXEL : FG_ALGO_GRAIN;
    fd.inv_e_pi_r2 = d.inv_e_pi_r2;
    const size_t nout = d.output_width * d.output_height, nr = p.color_mode == ColorMode::Rgb ? 3 : 1;
    RgbImage img;
    img.width = (uint32_t)d.output_width;
    img.height = (uint32_t)d.output_height;
    img.data.resize(nout * 3);
    std::vector<float> planes(planes_out ? nr * nout : 0);
    CancelThunk thunk{cancel};
    const std::vector<float> &offs = algo == Algo::Pixel ? d.offsets_input : d.offsets;
    const int rc = fg_render_frame(ctx, &fd, rgb, offs.data(), img.data.data(), planes_out ? planes.data() : nullptr,
                                   cancel ? &CancelThunk::call : nullptr, &thunk);
    if (rc == FG_ECANCELLED) throw RenderError(RenderError::Cancelled, "cancelled");
    if (rc != FG_OK) {
        const std::string msg = std::string("gpu error: ") + fg_last_error(ctx);
        if (rc == FG_ENOMEM || rc == FG_EHIP) invalidate_context(ctx);
        throw RenderError(RenderError::Gpu, msg);
    }
    if (planes_out)
        for (size_t i = 0; i < nr; ++i)
            planes_out->emplace_back(d.output_width, d.output_height,
                                     std::vector<float>(planes.begin() + i * nout, planes.begin() + (i + 1) * nout));
    check_cancel(cancel);
    return {std::move(img), make_stats(p, d, algo)};
}

// File input renders the whole frame on the device (the same image as the
// per-plane route through render_workspace, one upload and one download).
std::pair<RgbImage, RenderStats> render_to_image(const Params &p)
{
    size_t w = 0, h = 0;
    const std::vector<float> rgb = load_image_rgb_f32(p.input_path, &w, &h);
    return render_rgb_frame(rgb.data(), w, h, p, nullptr, nullptr);
}

RenderStats render(const Params &p)
{
    auto r = render_to_image(p);
    const std::string format = resolve_format(p);
    save_image(r.first, p.output_path, format);
    return r.second;
}

std::pair<RgbImage, RenderStats> render_with_input_image(const InputImage &input, const Params &p)
{
    return render_workspace(input.to_workspace(), p, nullptr, nullptr);
}

std::pair<RgbImage, RenderStats> render_with_input_image_cancelable(const InputImage &input, const Params &p,
                                                                    const CancelCheck &cancel)
{
    return render_workspace(input.to_workspace(), p, &cancel, nullptr);
}

RenderStats dry_run(const Params &p)
{
    const InputImage in = InputImage::from_path(p.input_path, p.color_mode, p.roi);
    return dry_run_with_input_image(in, p);
}

RenderStats dry_run_with_input_image(const InputImage &input, const Params &p)
{
    auto dv = derive_for_workspace(input.ws, p, nullptr);
    return make_stats(p, dv.first, dv.second);
}

RenderStats dry_run_with_input_image_cancelable(const InputImage &input, const Params &p,
                                                const CancelCheck &cancel)
{
    auto dv = derive_for_workspace(input.ws, p, &cancel);
    return make_stats(p, dv.first, dv.second);
}

}  // namespace film_grain
