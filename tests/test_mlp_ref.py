"""The CPU reference of the error model is itself pinned: with the BF16
rounding switched off, its analytic gradient equals float64 central finite
differences of its loss, and its forward pass equals a plain restatement of
the network.  (A dropped term, wrong sign or transposed operand in the
reference's backward pass fails the first check.)"""
import numpy as np
import pytest

from tests import mlp_ref


def _no_rounding(monkeypatch):
    monkeypatch.setattr(mlp_ref, "bf16", lambda x: np.asarray(x, dtype=np.float64))


@pytest.mark.parametrize("kind", ["squared", "nll"])
def test_gradient_equals_finite_differences(monkeypatch, kind):
    _no_rounding(monkeypatch)
    rng = np.random.default_rng(0)
    n = 40
    # smaller weights for the likelihood keep y = O(1), where exp(-2y) and
    # its finite differences stay well conditioned
    P = rng.normal(0, 0.4 if kind == "squared" else 0.15, mlp_ref.N_PARAMS)
    P[mlp_ref.P_W1 + 12 * 64:mlp_ref.P_B1] = 0.0
    X = np.zeros((n, 16))
    X[:, :12] = rng.uniform(-1, 1, (n, 12))
    t = rng.normal(0, 1, n)

    def loss(p):
        y = mlp_ref.forward(p, X)[-1]
        if kind == "nll":  # written out here, not through the reference
            return (y + 0.5 * t * t * np.exp(-2.0 * y)).sum() / n
        return ((y - t) ** 2).sum() / n

    g, lsum = mlp_ref.grad(P, X, t, loss=kind)
    assert abs(lsum / n - loss(P)) <= 1e-9 * abs(loss(P)) + 1e-12
    idx = np.concatenate([rng.choice(np.arange(mlp_ref.P_W1, mlp_ref.P_W1 + 12 * 64), 40, replace=False),
                          rng.choice(np.arange(mlp_ref.P_B1, mlp_ref.N_PARAMS), 80, replace=False),
                          [mlp_ref.P_B3]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(P)
        e[i] = h
        fd = (loss(P + e) - loss(P - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_forward_is_the_network(monkeypatch):
    _no_rounding(monkeypatch)
    rng = np.random.default_rng(1)
    P = rng.normal(0, 0.4, mlp_ref.N_PARAMS)
    X = rng.uniform(-1, 1, (7, 16))
    W1, b1, W2, b2, w3, b3 = mlp_ref.unpack(P)
    W1, b1, W2, b2, w3, b3 = (np.asarray(a, np.float64) for a in (W1, b1, W2, b2, w3, b3))
    want = np.maximum(np.maximum(X @ W1 + b1, 0) @ W2 + b2, 0) @ w3 + b3
    assert np.allclose(mlp_ref.forward(P, X)[-1], want, rtol=1e-6, atol=1e-6)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    got = mlp_ref.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0                  # halfway, even stays
    assert got[2] == 1.0 + 2 ** -6        # halfway, rounds to even (up)
    assert got[3] == 1.0
    assert got[4] == -2.5
    assert np.isfinite(got[5])
