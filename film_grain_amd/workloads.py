"""Synthetic workloads of the benchmark, the autotuner and the batch tests.

Each workload is a frame of synthetic intensity planes and the reference CLI
parameters used to render it (radius law, zoom, N, seed); `prepare` derives
the plane description, the lambda planes and the offsets through the host
library exactly as the reference's pipeline does before its per-plane
integrator call (src/lib.rs:126-165).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

WORKLOADS = {
    "pixel-c2": dict(w=1920, h=1080, rgb=True, radius=0.1, sd=0.05, dist="lognorm", zoom=2.0,
                     iters=256, seed=7, algo="pixel",
                     desc="pixel-wise 1920x1080 RGB -> 3840x2160, lognormal 0.1/0.05, sigma 0.8, N=256, seed 7"),
    "grain-c3": dict(w=4096, h=4096, rgb=False, radius=0.05, sd=0.0, dist="const", zoom=1.0,
                     iters=128, seed=11, algo="grain",
                     desc="grain-wise 4096x4096 luma ramp+5% noise, r=0.05 const, sigma 0.8, N=128, seed 11"),
    "pixel-c1": dict(w=256, h=256, rgb=False, radius=0.12, sd=0.0, dist="const", zoom=1.0,
                     iters=64, seed=42, algo="pixel",
                     desc="pixel-wise 256x256 luma, r=0.12 const, sigma 0.8, N=64, seed 42"),
}


def make_planes(name: str) -> np.ndarray:
    """(planes, h, w) f32 intensity planes of a workload."""
    from film_grain_amd import synth
    wl = WORKLOADS[name]
    if name == "pixel-c2":
        return synth.c2_rgb(wl["seed"], wl["w"], wl["h"])
    if name == "grain-c3":
        return synth.c3_ramp(wl["seed"], wl["w"])[None]
    return synth.c1_luma(wl["seed"], wl["w"])[None]


@dataclass
class Prepared:
    name: str
    algo: str
    planes: np.ndarray   # intensity planes
    lams: list           # lambda planes (host f32)
    offsets: np.ndarray  # offsets the integrator takes (input px for pixel-wise)
    derived: object      # host.DerivedC
    desc: object         # capi.PlaneDesc


@dataclass
class BatchImage:
    w: int
    h: int
    rgb: bool
    radius: float
    sd: float
    n_samples: int
    seed: int


def batch_images(master_seed: int = 2024, n_images: int = 64, min_side: int = 128,
                 max_side: int = 2048) -> list[BatchImage]:
    """The uneven batch: sides uniform in [min_side, max_side], radius
    U(0.05, 0.3) px, sigma_r / radius U(0, 0.5), N from {32, 64, 128, 256, 512},
    luma or RGB with equal odds, filter sigma 0.8, zoom 1."""
    rng = np.random.default_rng(master_seed)
    out = []
    for i in range(n_images):
        w = int(rng.integers(min_side, max_side + 1))
        h = int(rng.integers(min_side, max_side + 1))
        r = float(rng.uniform(0.05, 0.3))
        ratio = float(rng.uniform(0.0, 0.5))
        n = int(rng.choice([32, 64, 128, 256, 512]))
        rgb = bool(rng.integers(0, 2))
        out.append(BatchImage(w, h, rgb, r, ratio * r, n, master_seed * 1000 + i))
    return out


def prepare_batch(images: list[BatchImage]):
    """Per plane of every image: (plane description, lambda plane,
    offsets_input) for the pixel-wise path, with the image's derived
    parameters (sinusoid fields: 4 sinusoids + 2% noise per channel)."""
    from film_grain_amd import capi, host, synth
    planes = []
    for im in images:
        cli = host.CliArgs(radius=im.radius, radius_stddev=im.sd, radius_dist="lognorm" if im.sd > 0 else "const",
                           zoom=1.0, iters=im.n_samples, seed=im.seed, algo="pixel", sigma=0.8,
                           color_mode="rgb" if im.rgb else "luma")
        d, _offs, offs_in = host.derive(cli, im.w, im.h)
        desc = capi.make_desc(d.in_w, d.in_h, d.out_w, d.out_h, d.n_samples, d.dist_kind, d.seed, d.zoom,
                              d.delta, d.rm, d.radius_mean, d.log_mu, d.log_sigma)
        for c in range(3 if im.rgb else 1):
            u = synth.sinusoid_field(im.w, im.h, im.seed * 10 + c, n_waves=4, noise=0.02)
            nrm, _ = host.normalize_plane(u)
            planes.append((desc, host.lambda_plane(nrm, d.inv_e_pi_r2), np.ascontiguousarray(offs_in)))
    return planes


def prepare(name: str) -> Prepared:
    from film_grain_amd import capi, host
    wl = WORKLOADS[name]
    planes = make_planes(name)
    cli = host.CliArgs(radius=wl["radius"], radius_stddev=wl["sd"], radius_dist=wl["dist"], zoom=wl["zoom"],
                       iters=wl["iters"], seed=wl["seed"], algo=wl["algo"], sigma=0.8,
                       color_mode="rgb" if wl["rgb"] else "luma")
    d, offs, offs_in = host.derive(cli, wl["w"], wl["h"])
    lams = []
    for pl in planes:
        nrm, _ = host.normalize_plane(pl)
        lams.append(host.lambda_plane(nrm, d.inv_e_pi_r2))
    offsets = offs_in if wl["algo"] == "pixel" else offs
    desc = capi.make_desc(d.in_w, d.in_h, d.out_w, d.out_h, d.n_samples, d.dist_kind, d.seed, d.zoom,
                          d.delta, d.rm, d.radius_mean, d.log_mu, d.log_sigma)
    return Prepared(name, wl["algo"], planes, lams, np.ascontiguousarray(offsets), d, desc)
