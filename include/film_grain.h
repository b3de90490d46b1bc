/*
 * film_grain.h -- C ABI of the host library above the HIP integrators.
 *
 * The library restates the reference's host surface in C++ (film_grain_amd/
 * csrc/host/fg_host.h).  This C view exposes what a non-C++ caller needs:
 * parameter validation (ParamsBuilder::try_from + build, src/params.rs:141-221),
 * derived parameters (derive_common, src/model.rs:181-226), offsets
 * (make_offsets, src/rng.rs:9-24), the lambda pre-pass (normalize_plane /
 * lambda_plane, src/model.rs:228-265), routing (choose_algorithm,
 * src/choose.rs:4-26) and the render entry points (render,
 * render_with_input_image[_cancelable], dry_run, src/lib.rs:57-112).
 */
#ifndef FILM_GRAIN_H
#define FILM_GRAIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; one per RenderError variant (src/lib.rs:28-44). */
#define FGH_OK 0
#define FGH_EPARAMS 10      /* RenderError::Params */
#define FGH_EIMAGE 11       /* RenderError::Image */
#define FGH_EIO 12          /* RenderError::Io */
#define FGH_EGPU 13         /* RenderError::Gpu */
#define FGH_EUNSUPPORTED 14 /* RenderError::Unsupported (e.g. --device cpu) */
#define FGH_ECANCELLED 15   /* RenderError::Cancelled */
#define FGH_EMESSAGE 16     /* RenderError::Message */
#define FGH_ECAPACITY 17    /* caller's output buffer is too small */

/* CLI-level arguments (~ CliArgs, src/params.rs:93-114; the seed is 64-bit as
 * in Params and the viewer).  NULL strings mean "not given"; max_radius and
 * cell default to "quantile 0.999" and "approx mu_r" when NULL. */
typedef struct {
    const char *input_path;
    const char *output_path;
    uint32_t radius_dist; /* 0 const, 1 lognorm */
    double radius_mean, radius_stddev, zoom, sigma_px;
    uint32_t n_samples;
    uint32_t algo;        /* 0 auto, 1 grain, 2 pixel */
    const char *max_radius;
    const char *cell;
    uint32_t color_mode;  /* 0 luma, 1 rgb */
    uint32_t device;      /* 0 cpu, 1 gpu (alias of hip), 2 hip */
    const char *roi;
    const char *size;
    uint64_t seed;
    int32_t dry_run, explain;
    const char *output_format;
} fgh_cli_args;

/* Derived parameters and RenderStats (src/model.rs:167-179, src/lib.rs:46-55). */
typedef struct {
    uint32_t in_w, in_h, out_w, out_h, n_samples;
    uint32_t algorithm;   /* 1 grain, 2 pixel (after auto routing) */
    uint32_t dist_kind;
    uint64_t seed;
    float zoom, delta, rm, radius_mean, radius_stddev, log_mu, log_sigma, inv_e_pi_r2;
    float sigma_ratio, rm_ratio;
} fgh_derived;

typedef int (*fgh_cancel_fn)(void *user);

/* Validates args and derives the parameters for an input of in_w x in_h;
 * offsets_xy / offsets_input_xy (2*n_samples floats each) may be NULL. */
int fgh_derive(const fgh_cli_args *args, uint32_t in_w, uint32_t in_h, fgh_derived *out,
               float *offsets_xy, float *offsets_input_xy, char *err, size_t errlen);
int fgh_make_offsets(uint64_t seed, uint32_t n, float sigma, float *out_xy);
float fgh_normalize_plane(const float *src, size_t n, float *dst);
void fgh_lambda_plane(const float *normalized, size_t n, float inv_e_pi_r2, float *dst);
double fgh_inverse_normal_cdf(double p);

/* Renders an interleaved RGB f32 image (values as image::to_rgb32f gives
 * them) with the CLI-level args: ~ render_with_input_image_cancelable.
 * out_rgb8 receives out_w*out_h*3 bytes; out_planes (optional) receives the
 * rendered f32 planes (1 for luma, 3 for rgb), each out_w*out_h. */
int fgh_render_rgb(const fgh_cli_args *args, const float *rgb, uint32_t w, uint32_t h,
                   uint8_t *out_rgb8, size_t out_rgb8_cap, float *out_planes, size_t out_planes_cap,
                   fgh_derived *stats, fgh_cancel_fn cancel, void *cancel_user, char *err,
                   size_t errlen);
/* The same as fgh_render_rgb, bit for bit, with the whole frame on the device
 * (colour split, lambda pre-pass, integrator and merge behind one upload and
 * one download: fg_render_frame); the ROI crop stays on the host. */
int fgh_render_rgb_frame(const fgh_cli_args *args, const float *rgb, uint32_t w, uint32_t h, uint8_t *out_rgb8,
                         size_t out_rgb8_cap, float *out_planes, size_t out_planes_cap, fgh_derived *stats,
                         fgh_cancel_fn cancel, void *cancel_user, char *err, size_t errlen);
/* out[i] = logf(x[i]) of the C library (the reference's f32::ln). */
void fgh_logf_array(const float *x, size_t n, float *out);
/* Workspace::from_rgb_f32's planes: luma mode Y, Cb, Cr; rgb mode R, G, B. */
int fgh_color_split(const float *rgb, uint32_t w, uint32_t h, uint32_t color_mode, float *p0, float *p1, float *p2);
/* File to file (~ render, src/lib.rs:57-71). */
int fgh_render_file(const fgh_cli_args *args, fgh_derived *stats, char *err, size_t errlen);
/* ~ dry_run (src/lib.rs:95-98). */
int fgh_dry_run(const fgh_cli_args *args, fgh_derived *stats, char *err, size_t errlen);

/* Colour conversions alone: Workspace from an RGB f32 image (luma: BT.709
 * Y/Cb/Cr split; rgb: three planes) then into_rgb_image, with no rendering in
 * between (src/color.rs:66-114, 158-213).  out_rgb8: w*h*3 bytes. */
int fgh_color_roundtrip(const float *rgb, uint32_t w, uint32_t h, uint32_t color_mode,
                        uint8_t *out_rgb8);
/* Reads PNG / PGM / PPM / PFM as RGB f32 (image::to_rgb32f semantics).  With
 * out_rgb NULL only the size is returned. */
int fgh_load_image(const char *path, uint32_t *w, uint32_t *h, float *out_rgb, size_t cap, char *err,
                   size_t errlen);
/* Writes an RGB8 image as "png" or "ppm". */
int fgh_save_rgb8(const char *path, const uint8_t *rgb8, uint32_t w, uint32_t h, const char *format,
                  char *err, size_t errlen);

#ifdef __cplusplus
}
#endif
#endif
