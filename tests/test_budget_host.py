"""Host logic of the learned sample budget (film_grain_amd/budget.py) on the
CPU: the frame-scale search, the fixed-N interpolation, and the arithmetic
that makes the two-halves N = 4096 render exact."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from film_grain_amd import budget  # noqa: E402


def test_halves_mean_is_the_4096_render():
    # every pair of half counts: (c1 * 2^-11 + c2 * 2^-11) * 0.5 == (c1 + c2) * 2^-12 in f32
    c = np.arange(0, 2049, dtype=np.float32)
    inv_half = np.float32(1.0) / np.float32(2048)
    inv_full = np.float32(1.0) / np.float32(4096)
    a = (c * inv_half)[:, None]
    b = (c * inv_half)[None, :]
    got = (a + b) * np.float32(0.5)
    want = (c[:, None] + c[None, :]) * inv_full
    assert got.dtype == np.float32 and np.array_equal(got, want)


def _brute_scale(s, target, n_min, n_max):
    """Smallest c on a fine geometric grid meeting the target (for checking)."""
    cs = np.exp(np.linspace(math.log(1e-2), math.log(1e6), 20001))
    for c in cs:
        n = np.clip(np.ceil(c * s), n_min, n_max)
        if np.mean(s * s * (1 / n - 1 / budget.N_REF)) <= target:
            return c
    return cs[-1]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_choose_scale_meets_the_target_with_the_fewest_samples(seed):
    rng = np.random.default_rng(seed)
    s = np.sqrt(rng.uniform(0, 0.25, 5000)).astype(np.float32)
    s[:200] = 0.0
    for n_target in (64, 256, 900):
        target = float(np.mean(s.astype(np.float64) ** 2) * (1 / n_target - 1 / budget.N_REF))
        c = budget.choose_scale(torch.from_numpy(s), target, 16, 1024)
        n = np.clip(np.ceil(c * s.astype(np.float64)), 16, 1024)
        m = np.mean(s.astype(np.float64) ** 2 * (1 / n - 1 / budget.N_REF))
        assert m <= target * (1 + 1e-5)
        # not wasteful: the brute-force smallest scale is within 1%
        cb = _brute_scale(s.astype(np.float64), target, 16, 1024)
        assert c <= cb * 1.01 + 1e-9, (c, cb)
        # allocation proportional to s uses fewer samples than fixed N at equal error
        assert n.mean() <= n_target * 1.001


def test_choose_scale_unreachable_target_gives_the_cap():
    s = torch.full((1000,), 0.5)
    target = 0.25 * (1 / 2048 - 1 / budget.N_REF)  # below what n = 1024 reaches
    c = budget.choose_scale(s, target, 16, 1024)
    assert math.ceil(c * 0.5) >= 1024


def test_ms_at_rmse_interpolates_log_log():
    curve = [(64, 10.0, 0.04), (256, 30.0, 0.02), (1024, 100.0, 0.01)]
    assert budget.ms_at_rmse(curve, 0.02) == pytest.approx(30.0)
    mid = budget.ms_at_rmse(curve, math.sqrt(0.02 * 0.01))
    assert mid == pytest.approx(math.sqrt(30.0 * 100.0))
    assert budget.ms_at_rmse(curve, 0.005) is None
    assert budget.ms_at_rmse(curve, 0.05) is None


def test_s_ratio_bounds():
    r = torch.full((100,), 0.3)
    assert budget.s_ratio([r]) == pytest.approx(1.0)
    r2 = torch.cat([torch.full((50,), 0.5), torch.zeros(50)])
    assert budget.s_ratio([r2]) == pytest.approx(2.0)


def test_init_params_layout():
    from film_grain_amd import capi
    P = budget.init_params(0)
    assert P.shape == (capi.MLP_PARAMS,)
    W1 = P[:1024].reshape(16, 64)
    assert np.all(W1[12:] == 0) and np.all(P[1024:1088] == 0)
    assert P[5312] == pytest.approx(math.log(0.5 / budget.SIGMA64_TO_S))


def test_scale_for_mean_budget_spends_at_most_the_total():
    rng = np.random.default_rng(3)
    s = torch.from_numpy(np.sqrt(rng.uniform(0, 0.25, 20000)).astype(np.float32))
    for mb in (20.0, 64.0, 300.0, 900.0):
        c = budget.scale_for_mean_budget(s, mb, 16, 1024)
        n = torch.clamp(torch.ceil(c * s.double()), 16, 1024)
        assert float(n.mean()) <= mb * (1 + 1e-6)
        n2 = torch.clamp(torch.ceil(c * 1.02 * s.double()), 16, 1024)
        assert float(n2.mean()) > mb * 0.99  # close to the largest such scale
