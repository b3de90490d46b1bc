"""Seeded synthetic intensity fields for the benchmark workloads and tests.

The reference's own bench tool draws constant / step / ramp / natural fields
(tools/film_grain_bench.py:147-177); the workloads here are the sinusoid-plus-
noise fields and the noisy ramp that the benchmark configurations name.  Fields
are f32 in [0, 0.999], row-major (h, w), and deterministic for a given seed
(numpy PCG64).
"""
from __future__ import annotations

import numpy as np


def sinusoid_field(w: int, h: int, seed: int, n_waves: int = 4, noise: float = 0.02,
                   flat_fraction: float = 0.0, flat_values=(0.0, 0.98)) -> np.ndarray:
    """Sum of ``n_waves`` random 2-D sinusoids (1-8 cycles per image) plus
    uniform noise of half-width ``noise``; optionally ``flat_fraction`` of the
    area replaced by axis-aligned flat patches (half at each of
    ``flat_values``); clamped to [0, 0.999]."""
    rng = np.random.default_rng(seed)
    fx = rng.integers(1, 9, size=n_waves)
    fy = rng.integers(1, 9, size=n_waves)
    amp = rng.uniform(0.5, 1.0, size=n_waves)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=n_waves)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    acc = np.zeros((h, w), dtype=np.float64)
    for k in range(n_waves):
        acc += amp[k] * np.sin(2.0 * np.pi * (fx[k] * x / w + fy[k] * y / h) + phase[k])
    field = 0.5 + 0.5 * acc / amp.sum()
    field += rng.uniform(-noise, noise, size=(h, w))
    if flat_fraction > 0.0:
        target = flat_fraction * w * h
        covered = np.zeros((h, w), dtype=bool)
        n = 0
        while covered.sum() < target and n < 10000:
            pw = int(rng.integers(max(2, w // 20), max(3, w // 6)))
            ph = int(rng.integers(max(2, h // 20), max(3, h // 6)))
            px = int(rng.integers(0, max(1, w - pw)))
            py = int(rng.integers(0, max(1, h - ph)))
            field[py:py + ph, px:px + pw] = flat_values[n % 2]
            covered[py:py + ph, px:px + pw] = True
            n += 1
    return np.clip(field, 0.0, 0.999).astype(np.float32)


def ramp_field(w: int, h: int, seed: int, noise: float = 0.05) -> np.ndarray:
    """u = 0.999 * x / (w - 1) plus uniform noise of half-width ``noise``,
    clamped to [0, 0.999]."""
    rng = np.random.default_rng(seed)
    x = np.arange(w, dtype=np.float64)
    base = 0.999 * x / max(w - 1, 1)
    field = np.broadcast_to(base, (h, w)) + rng.uniform(-noise, noise, size=(h, w))
    return np.clip(field, 0.0, 0.999).astype(np.float32)


def c1_luma(seed: int = 42, size: int = 256) -> np.ndarray:
    """Plain pixel-wise CPU workload: 256x256 luma, 4 sinusoids + 2% noise."""
    return sinusoid_field(size, size, seed, n_waves=4, noise=0.02)


def c2_rgb(seed: int = 7, w: int = 1920, h: int = 1080) -> np.ndarray:
    """Pixel-wise 1080p RGB workload: per channel 8 sinusoids + 2% noise with
    10% flat patches at 0.0 and 0.98.  Returns (3, h, w)."""
    return np.stack([sinusoid_field(w, h, seed * 1000 + c, n_waves=8, noise=0.02,
                                    flat_fraction=0.10) for c in range(3)])


def c3_ramp(seed: int = 11, size: int = 4096) -> np.ndarray:
    """Grain-wise workload: 4096x4096 luma ramp + 5% noise."""
    return ramp_field(size, size, seed, noise=0.05)
