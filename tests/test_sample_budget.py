"""Adaptive sampling rests on the offsets of a smaller sample count being a
prefix of a larger count's (the offsets are one seeded Gaussian stream,
src/rng.rs:9-24), and on the derived parameters other than the sample count
not depending on it (src/derive.rs)."""
import numpy as np
import pytest


@pytest.mark.parametrize("seed", [0, 7, 42, 2024])
def test_offsets_prefix(oracle, seed):
    full = oracle.make_offsets(seed, 2048, 0.8)
    for n in [1, 2, 15, 16, 37, 255, 256, 1000, 1024]:
        assert np.array_equal(oracle.make_offsets(seed, n, 0.8), full[:n]), n


def test_derived_independent_of_sample_count(oracle):
    a = oracle.derive(0.1, 0.05, "lognorm", 2.0, 0.8, 256, seed=7, in_w=64, in_h=48)
    b = oracle.derive(0.1, 0.05, "lognorm", 2.0, 0.8, 16, seed=7, in_w=64, in_h=48)
    for f in ("delta", "rm", "log_mu", "log_sigma", "inv_e_pi_r2", "out_w", "out_h"):
        assert getattr(a, f) == getattr(b, f), f
