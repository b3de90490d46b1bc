// fg_image_io.cpp -- image files for the library and CLI: PNG (8/16-bit grey,
// grey+alpha, RGB, RGBA, 8-bit palette; non-interlaced) through zlib, binary
// PGM/PPM (8/16-bit) and PFM.  Samples become RGB f32 as image::to_rgb32f does
// (v / 255 or v / 65535; alpha dropped).  Output: PNG or PPM of an RgbImage.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

#include "fg_host.h"

namespace film_grain {

namespace {

std::vector<uint8_t> read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw RenderError(RenderError::Io, "io error: cannot open '" + path + "'");
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

[[noreturn]] void bad(const std::string &m) { throw RenderError(RenderError::Image, "image error: " + m); }

std::vector<float> decode_png(const std::vector<uint8_t> &f, size_t *w_out, size_t *h_out)
{
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) bad("not a PNG file");
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, palette;
    while (pos + 12 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        if (pos + 12 + (size_t)len > f.size()) bad("truncated PNG chunk");
        const char *type = reinterpret_cast<const char *>(&f[pos + 4]);
        const uint8_t *data = &f[pos + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) bad("bad IHDR");
            w = be32(data);
            h = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            palette.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    if (w == 0 || h == 0) bad("PNG has no image");
    if (interlace) bad("interlaced PNG is not supported");
    int channels;
    switch (ctype) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 3: channels = 1; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: bad("unsupported PNG colour type");
    }
    if (!(depth == 8 || depth == 16) || (ctype == 3 && depth != 8)) bad("unsupported PNG bit depth");
    const size_t bpp = (size_t)channels * (depth / 8);
    const size_t stride = (size_t)w * bpp;
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf rawlen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) != Z_OK || rawlen != raw.size())
        bad("corrupt PNG data");
    std::vector<uint8_t> img(stride * h);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t *src = &raw[y * (stride + 1) + 1];
        uint8_t *dst = &img[y * stride];
        const uint8_t *up = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? dst[i - bpp] : 0, b = up ? up[i] : 0,
                      c = (up && i >= bpp) ? up[i - bpp] : 0;
            int v = src[i];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: {
                const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
                v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                break;
            }
            default: bad("bad PNG filter");
            }
            dst[i] = (uint8_t)v;
        }
    }
    std::vector<float> rgb((size_t)w * h * 3);
    const float maxv = depth == 16 ? 65535.0f : 255.0f;
    for (size_t k = 0; k < (size_t)w * h; ++k) {
        float s[4];
        for (int c = 0; c < channels; ++c) {
            const uint8_t *p = &img[k * bpp + (size_t)c * (depth / 8)];
            s[c] = depth == 16 ? (float)((p[0] << 8) | p[1]) : (float)p[0];
        }
        if (ctype == 3) {
            const size_t idx = (size_t)s[0];
            if (3 * idx + 2 >= palette.size()) bad("palette index out of range");
            for (int c = 0; c < 3; ++c) rgb[3 * k + c] = (float)palette[3 * idx + c] / 255.0f;
        } else if (channels <= 2) {
            rgb[3 * k] = rgb[3 * k + 1] = rgb[3 * k + 2] = s[0] / maxv;
        } else {
            for (int c = 0; c < 3; ++c) rgb[3 * k + c] = s[c] / maxv;
        }
    }
    *w_out = w;
    *h_out = h;
    return rgb;
}

size_t pnm_token(const std::vector<uint8_t> &f, size_t &pos)
{
    for (;;) {
        while (pos < f.size() && std::isspace(f[pos])) ++pos;
        if (pos < f.size() && f[pos] == '#') {
            while (pos < f.size() && f[pos] != '\n') ++pos;
            continue;
        }
        break;
    }
    size_t v = 0;
    bool any = false;
    while (pos < f.size() && std::isdigit(f[pos])) {
        v = v * 10 + (f[pos++] - '0');
        any = true;
    }
    if (!any) bad("malformed PNM header");
    return v;
}

std::vector<float> decode_pnm(const std::vector<uint8_t> &f, size_t *w, size_t *h)
{
    const bool rgb = f[1] == '6';
    size_t pos = 2;
    *w = pnm_token(f, pos);
    *h = pnm_token(f, pos);
    const size_t maxv = pnm_token(f, pos);
    ++pos;  // single whitespace
    if (*w == 0 || *h == 0 || maxv == 0 || maxv > 65535) bad("bad PNM header");
    const size_t bps = maxv > 255 ? 2 : 1, ch = rgb ? 3 : 1;
    if (f.size() < pos + (*w) * (*h) * ch * bps) bad("truncated PNM data");
    std::vector<float> out((*w) * (*h) * 3);
    for (size_t k = 0; k < (*w) * (*h); ++k) {
        for (size_t c = 0; c < 3; ++c) {
            const size_t src = (k * ch + (rgb ? c : 0)) * bps + pos;
            const float v = bps == 2 ? (float)((f[src] << 8) | f[src + 1]) : (float)f[src];
            out[3 * k + c] = v / (float)maxv;
        }
    }
    return out;
}

std::vector<float> decode_pfm(const std::vector<uint8_t> &f, size_t *w, size_t *h)
{
    const bool rgb = f[1] == 'F';
    size_t pos = 2;
    *w = pnm_token(f, pos);
    *h = pnm_token(f, pos);
    while (pos < f.size() && std::isspace(f[pos])) ++pos;
    size_t e = pos;
    while (e < f.size() && !std::isspace(f[e])) ++e;
    const double scale = std::atof(std::string(f.begin() + pos, f.begin() + e).c_str());
    pos = e + 1;
    const size_t ch = rgb ? 3 : 1;
    if (*w == 0 || *h == 0 || f.size() < pos + (*w) * (*h) * ch * 4) bad("truncated PFM data");
    const bool little = scale < 0;
    std::vector<float> out((*w) * (*h) * 3);
    for (size_t y = 0; y < *h; ++y) {
        const size_t row = (*h - 1 - y);  // PFM rows run bottom to top
        for (size_t x = 0; x < *w; ++x)
            for (size_t c = 0; c < 3; ++c) {
                const uint8_t *p = &f[pos + ((row * (*w) + x) * ch + (rgb ? c : 0)) * 4];
                uint32_t b = little ? (uint32_t)p[0] | p[1] << 8 | p[2] << 16 | (uint32_t)p[3] << 24
                                    : (uint32_t)p[3] | p[2] << 8 | p[1] << 16 | (uint32_t)p[0] << 24;
                float v;
                std::memcpy(&v, &b, 4);
                out[3 * (y * (*w) + x) + c] = v;
            }
    }
    return out;
}

void put32(std::vector<uint8_t> &o, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}

void chunk(std::vector<uint8_t> &o, const char *type, const std::vector<uint8_t> &data)
{
    put32(o, (uint32_t)data.size());
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put32(o, (uint32_t)crc32(0, &o[start], (uInt)(o.size() - start)));
}

std::vector<uint8_t> encode_png(const RgbImage &img)
{
    std::vector<uint8_t> o = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    std::vector<uint8_t> ihdr;
    put32(ihdr, img.width);
    put32(ihdr, img.height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk(o, "IHDR", ihdr);
    const size_t stride = (size_t)img.width * 3;
    std::vector<uint8_t> raw((stride + 1) * img.height);
    for (uint32_t y = 0; y < img.height; ++y) {
        raw[y * (stride + 1)] = 0;
        std::memcpy(&raw[y * (stride + 1) + 1], &img.data[y * stride], stride);
    }
    uLongf clen = compressBound((uLong)raw.size());
    std::vector<uint8_t> comp(clen);
    if (compress2(comp.data(), &clen, raw.data(), (uLong)raw.size(), 6) != Z_OK) bad("PNG encode failed");
    comp.resize(clen);
    chunk(o, "IDAT", comp);
    chunk(o, "IEND", {});
    return o;
}

}  // namespace

std::vector<float> load_image_rgb_f32(const std::string &path, size_t *w, size_t *h)
{
    const std::vector<uint8_t> f = read_file(path);
    if (f.size() >= 8 && f[0] == 0x89 && f[1] == 'P') return decode_png(f, w, h);
    if (f.size() >= 2 && f[0] == 'P' && (f[1] == '5' || f[1] == '6')) return decode_pnm(f, w, h);
    if (f.size() >= 2 && f[0] == 'P' && (f[1] == 'F' || f[1] == 'f')) return decode_pfm(f, w, h);
    bad("unsupported image format (PNG, PGM/PPM and PFM are supported)");
}

void save_image(const RgbImage &img, const std::string &path, const std::string &format)
{
    std::vector<uint8_t> bytes;
    if (format == "png") {
        bytes = encode_png(img);
    } else if (format == "ppm" || format == "pnm") {
        const std::string hdr = "P6\n" + std::to_string(img.width) + " " + std::to_string(img.height) + "\n255\n";
        bytes.assign(hdr.begin(), hdr.end());
        bytes.insert(bytes.end(), img.data.begin(), img.data.end());
    } else {
        throw RenderError(RenderError::Unsupported, "unsupported: image format '" + format +
                                                        "' is not available in this build (png, ppm)");
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) throw RenderError(RenderError::Io, "io error: cannot write '" + path + "'");
    f.write(reinterpret_cast<const char *>(bytes.data()), (std::streamsize)bytes.size());
}

// src/lib.rs:188-203: explicit format, else the output extension, else PNG.
std::string resolve_format(const Params &p)
{
    std::string token;
    if (p.output_format) {
        token = *p.output_format;
    } else {
        const size_t slash = p.output_path.find_last_of('/');
        const size_t dot = p.output_path.find_last_of('.');
        if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "png";
        token = p.output_path.substr(dot + 1);
    }
    size_t b = 0, e = token.size();
    while (b < e && std::isspace((unsigned char)token[b])) ++b;
    while (e > b && std::isspace((unsigned char)token[e - 1])) --e;
    token = token.substr(b, e - b);
    while (!token.empty() && token[0] == '.') token.erase(0, 1);
    std::string l = token;
    for (auto &c : l) c = (char)std::tolower((unsigned char)c);
    static const char *known[] = {"png", "jpg", "jpeg", "gif", "webp", "tif", "tiff", "tga", "dds",
                                  "bmp", "ico", "hdr", "exr", "pbm", "pam", "ppm", "pgm", "pnm",
                                  "ff", "farbfeld", "avif", "qoi"};
    for (const char *k : known)
        if (l == k) return l == "jpeg" ? "jpg" : (l == "tiff" ? "tif" : l);
    throw RenderError(RenderError::Message, "unsupported or unknown image format '" + token + "'");
}

}  // namespace film_grain
