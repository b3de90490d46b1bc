"""Shared builders for parity cases: a seeded intensity plane, the oracle's
derived parameters, offsets and lambda plane, and the matching C-ABI plane
description.  The workloads mirror the benchmark configurations at sizes the
CPU oracle finishes in seconds."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from film_grain_amd import synth


@dataclass
class Case:
    name: str
    u: np.ndarray             # intensity plane (h, w) f32
    lam: np.ndarray           # lambda plane (oracle lambda_plane)
    offsets: np.ndarray       # Derived::offsets (output px), (N, 2) f32
    offsets_input: np.ndarray  # Derived::offsets_input (input px)
    derived: object           # oracle FgoDerived
    plane: object             # oracle FgoPlane
    desc: object              # capi PlaneDesc
    n_samples: int
    seed: int
    zoom: float
    dist: str
    radius: float


def build_case(O, name, u, radius, stddev=0.0, dist="const", zoom=1.0, sigma=0.8, n_samples=64,
               seed=42, cell_delta=None, max_radius=("quantile", 0.999)) -> Case:
    from film_grain_amd import capi
    u = np.ascontiguousarray(u, dtype=np.float32)
    h, w = u.shape
    d = O.derive(radius, stddev, dist, zoom, sigma, n_samples, max_radius=max_radius,
                 cell_delta=cell_delta, seed=seed, in_w=w, in_h=h)
    norm, _ = O.normalize_plane(u)
    lam = O.lambda_plane(norm, d.inv_e_pi_r2)
    offs = O.make_offsets(seed, n_samples, sigma)
    offs_in = (offs / np.float32(zoom)).astype(np.float32)
    p = O.plane(d, n_samples, seed, zoom, dist, radius)
    desc = capi.make_desc(d.in_w, d.in_h, d.out_w, d.out_h, n_samples, 1 if dist == "lognorm" else 0,
                          seed, zoom, d.delta, d.rm, radius, d.log_mu, d.log_sigma)
    return Case(name, u, lam, offs, offs_in, d, p, desc, n_samples, seed, zoom, dist, radius)


def c1_case(O, size=64, n_samples=64):
    """Plain pixel-wise: r = 0.12 const, sigma 0.8, zoom 1, seed 42."""
    return build_case(O, f"c1_{size}", synth.c1_luma(42, size), 0.12, n_samples=n_samples, seed=42)


def c2_case(O, w=96, h=64, n_samples=256, channel=0):
    """Pixel-wise lognormal: mean 0.1 / sd 0.05, sigma 0.8, N=256, zoom 2, seed 7."""
    u = synth.c2_rgb(7, w, h)[channel]
    return build_case(O, f"c2_{w}x{h}_c{channel}", u, 0.1, 0.05, "lognorm", zoom=2.0,
                      n_samples=n_samples, seed=7)


def c3_case(O, size=96, n_samples=128):
    """Grain-wise: r = 0.05 const, ramp + 5% noise, N=128, zoom 1, seed 11."""
    return build_case(O, f"c3_{size}", synth.c3_ramp(11, size), 0.05, n_samples=n_samples, seed=11)


def tiles(rng, out_w, out_h, tile, count):
    """Random tile origins (x0, y0, x1, y1) inside the output."""
    out = []
    for _ in range(count):
        x0 = int(rng.integers(0, max(1, out_w - tile + 1)))
        y0 = int(rng.integers(0, max(1, out_h - tile + 1)))
        out.append((x0, y0, min(x0 + tile, out_w), min(y0 + tile, out_h)))
    return out


def oracle_plane(O, desc, n_samples=None):
    """The oracle's FgoPlane for a C-ABI plane description (optionally with
    another sample count, whose offsets are the first n of the same set)."""
    p = O.FgoPlane()
    p.in_w, p.in_h, p.out_w, p.out_h = desc.in_w, desc.in_h, desc.out_w, desc.out_h
    p.n_samples = desc.n_samples if n_samples is None else n_samples
    p.dist_kind, p.seed = desc.dist_kind, desc.seed
    p.zoom, p.delta, p.rm, p.radius_mean = desc.zoom, desc.delta, desc.rm, desc.radius_mean
    p.log_mu, p.log_sigma = desc.log_mu, desc.log_sigma
    return p


def oracle_frame(O, rgb, color_mode="luma", radius=0.1, radius_stddev=0.0, radius_dist="const", zoom=1.0,
                 sigma=0.8, iters=32, seed=5489, algo="pixel"):
    """A whole frame from the oracle alone: colour split (src/color.rs:158-213),
    normalize_plane + lambda_plane and the integrator per rendered plane, and
    the merge with the 8-bit quantisation (src/color.rs:66-114).  Returns
    (rgb8 (h, w, 3), rendered planes (n, h, w))."""
    mode = O.COLOR_RGB if color_mode == "rgb" else O.COLOR_LUMA
    planes = O.color_split(rgb, mode)
    outs = []
    for i in range(3 if mode == O.COLOR_RGB else 1):
        c = build_case(O, f"frame{i}", planes[i], radius, radius_stddev, radius_dist, zoom=zoom, sigma=sigma,
                       n_samples=iters, seed=seed)
        if algo == "pixel":
            outs.append(O.render_pixelwise(c.plane, c.lam, c.offsets_input))
        else:
            outs.append(O.render_grainwise(c.plane, c.lam, c.offsets))
    if mode == O.COLOR_RGB:
        rgb8 = O.color_merge(mode, outs[0], outs[1], outs[2])
    else:
        rgb8 = O.color_merge(mode, outs[0], planes[1], planes[2])
    return rgb8, np.stack(outs)
