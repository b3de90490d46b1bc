// fg_color.cpp -- colour workspaces (src/color.rs): luma mode renders Y and
// carries BT.709 Cb/Cr through; RGB mode renders the three planes
// independently.  Output is 8-bit with to_u8 = floor(255 v + 0.5), no gamma.
#include <algorithm>
#include <cmath>

#include "fg_host.h"

namespace film_grain {

namespace {
constexpr float kYR = 0.2126f, kYG = 0.7152f, kYB = 0.0722f;  // src/color.rs:9-11
constexpr float kCbDenom = 1.8556f, kCrDenom = 1.5748f;       // src/color.rs:12-13

inline float clamp01(float v)  // f32::clamp(0, 1) keeps NaN
{
    if (v < 0.0f) return 0.0f;
    if (v > 1.0f) return 1.0f;
    return v;
}

inline uint8_t to_u8(float v)  // src/color.rs:237-239; `as u8` saturates, NaN -> 0
{
    const float f = std::floor(clamp01(v) * 255.0f + 0.5f);
    if (!(f > 0.0f)) return 0;
    if (f >= 255.0f) return 255;
    return (uint8_t)f;
}
}  // namespace

// src/color.rs:158-213 (load_rgb_workspace / load_luma_workspace)
Workspace Workspace::from_rgb_f32(const float *rgb, size_t w, size_t h, ColorMode mode)
{
    Workspace ws;
    if (mode == ColorMode::Rgb) {
        ws.kind = Rgb;
        for (auto &p : ws.planes) p = Plane(w, h);
        for (size_t k = 0; k < w * h; ++k) {
            ws.planes[0].data[k] = clamp01(rgb[3 * k]);
            ws.planes[1].data[k] = clamp01(rgb[3 * k + 1]);
            ws.planes[2].data[k] = clamp01(rgb[3 * k + 2]);
        }
        return ws;
    }
    ws.kind = Luma;
    ws.y = Plane(w, h);
    ws.cb = Plane(w, h);
    ws.cr = Plane(w, h);
    for (size_t k = 0; k < w * h; ++k) {
        const float r = clamp01(rgb[3 * k]), g = clamp01(rgb[3 * k + 1]), b = clamp01(rgb[3 * k + 2]);
        const float luma = kYR * r + kYG * g + kYB * b;
        ws.cb.data[k] = (b - luma) / kCbDenom;
        ws.cr.data[k] = (r - luma) / kCrDenom;
        ws.y.data[k] = clamp01(luma);
    }
    return ws;
}

std::pair<size_t, size_t> Workspace::dimensions() const
{
    if (kind == Luma) return {y.width, y.height};
    return {planes[0].width, planes[0].height};
}

// src/color.rs:47-64: planes are processed one after another.
void Workspace::for_each_plane(const std::function<Plane(const Plane &, size_t)> &f)
{
    if (kind == Luma) {
        y = f(y, 0);
        return;
    }
    for (size_t i = 0; i < 3; ++i) planes[i] = f(planes[i], i);
}

// src/color.rs:66-114
RgbImage Workspace::into_rgb_image() const
{
    RgbImage img;
    if (kind == Luma) {
        const size_t w = y.width, h = y.height;
        const Plane cbp = (cb.width != w || cb.height != h) ? cb.resize_nearest(w, h) : cb;
        const Plane crp = (cr.width != w || cr.height != h) ? cr.resize_nearest(w, h) : cr;
        img.width = (uint32_t)w;
        img.height = (uint32_t)h;
        img.data.resize(w * h * 3);
        for (size_t k = 0; k < w * h; ++k) {
            const float yv = y.data[k], cbv = cbp.data[k], crv = crp.data[k];
            const float r = clamp01(yv + kCrDenom * crv);
            const float b = clamp01(yv + kCbDenom * cbv);
            const float g = clamp01((yv - kYR * r - kYB * b) / kYG);
            img.data[3 * k] = to_u8(r);
            img.data[3 * k + 1] = to_u8(g);
            img.data[3 * k + 2] = to_u8(b);
        }
        return img;
    }
    const size_t w = planes[0].width, h = planes[0].height;
    img.width = (uint32_t)w;
    img.height = (uint32_t)h;
    img.data.resize(w * h * 3);
    for (size_t k = 0; k < w * h; ++k) {
        img.data[3 * k] = to_u8(planes[0].data[k]);
        img.data[3 * k + 1] = to_u8(planes[1].data[k]);
        img.data[3 * k + 2] = to_u8(planes[2].data[k]);
    }
    return img;
}

// src/color.rs:215-231 (apply_roi) + from_dynamic_image
InputImage InputImage::from_rgb_f32(const float *rgb, size_t w, size_t h, ColorMode mode,
                                    const std::optional<Roi> &roi)
{
    InputImage in;
    in.color_mode = mode;
    if (!roi) {
        in.ws = Workspace::from_rgb_f32(rgb, w, h, mode);
        return in;
    }
    if (roi->x1 > w || roi->y1 > h) throw RenderError(RenderError::Message, "ROI exceeds image bounds");
    const size_t cw = roi->x1 - roi->x0, ch = roi->y1 - roi->y0;
    if (cw == 0 || ch == 0) throw RenderError(RenderError::Message, "ROI width and height must be positive");
    std::vector<float> crop(cw * ch * 3);
    for (size_t yy = 0; yy < ch; ++yy)
        std::copy(rgb + ((roi->y0 + yy) * w + roi->x0) * 3, rgb + ((roi->y0 + yy) * w + roi->x0 + cw) * 3,
                  crop.begin() + yy * cw * 3);
    in.ws = Workspace::from_rgb_f32(crop.data(), cw, ch, mode);
    return in;
}

InputImage InputImage::from_path(const std::string &path, ColorMode mode, const std::optional<Roi> &roi)
{
    size_t w = 0, h = 0;
    const std::vector<float> rgb = load_image_rgb_f32(path, &w, &h);
    return from_rgb_f32(rgb.data(), w, h, mode, roi);
}

}  // namespace film_grain
