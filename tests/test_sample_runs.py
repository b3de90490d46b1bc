"""The cyclic-run query of the sorted sample order is a superset.

Both splat kernels (fg_pixelwise_hbm.hip, fg_grainwise.hip) walk, per grain,
only the samples whose offset fraction frac(X) lies in a widened cyclic
interval, looked up through a bucket index.  A sample outside that run must
never have an f32 candidate box with a pixel column.  This test evaluates the
box and the run in IEEE f32 (numpy, no fused operations) on random and
adversarial grain positions (centres at integers and half-integers, +-1e-4)
and checks that no sample with a non-empty box is left out.  It also checks
the run stays within a small factor of the true candidate count.
"""
import numpy as np
import pytest

f32 = np.float32


def frac(x):
    f = (x - np.floor(x)).astype(f32)
    return np.where(f < f32(1), f, f32(0)).astype(f32)


def run_positions(A, rz, eps, qidx, Q, N):
    """The kernels' run query for one grain: sorted positions walked."""
    lo = f32(f32(A - f32(rz + rz)) - eps)
    hi = f32(A + eps)
    Qf = f32(Q)
    if not (lo * Qf > -2 * Qf) or not (hi * Qf < 2 * Qf):
        return np.arange(N)
    lq, hq = int(np.floor(lo * Qf)), int(np.floor(hi * Qf)) + 1
    if hq - lq >= Q:
        return np.arange(N)
    if lq < 0:
        lq += Q
        hq += Q
    lq = min(max(lq, 0), Q - 1)
    hq = min(max(hq, lq + 1), 2 * Q)
    p0 = qidx[lq]
    end = qidx[hq] if hq <= Q else N + qidx[hq - Q]
    return (p0 + np.arange(end - p0)) % N


@pytest.mark.parametrize("N,z,sd", [(7, 3.0, 0.5), (64, 1.0, 0.8), (256, 2.0, 0.4), (1000, 1.7, 1.5)])
def test_run_contains_every_candidate(N, z, sd):
    rng = np.random.default_rng(N)
    o = rng.normal(0, sd, size=N).astype(f32)
    X = (o * f32(z)).astype(f32)
    f = frac(X)
    order = np.lexsort((np.arange(N), f))
    Xs = X[order]
    Q = 64
    while Q < N and Q < 4096:
        Q <<= 1
    qidx = np.searchsorted(f[order], (np.arange(Q + 1) / Q).astype(f32), side="left")
    out_w = 3840
    c_out = (4000 + 10) * float(z) + 8
    eps = f32(c_out * 2.0 ** -20 + 2.0 ** -16)
    M = 2000
    bx = (rng.uniform(-50, 4000, M).astype(f32) * f32(z) - f32(0.5)).astype(f32)
    bx[: M // 2] = (np.round(bx[: M // 2] * 2) / 2 + rng.choice([-1e-4, 0, 1e-4], M // 2)).astype(f32)
    rz = (rng.uniform(0.001, 0.8, M).astype(f32) * f32(z) + f32(4e-3 + 2e-6 * out_w)).astype(f32)
    walked = true = 0
    for b, r in zip(bx, rz):
        cxk = (b + Xs).astype(f32)
        has_col = np.ceil((cxk - r).astype(f32)) <= np.floor((cxk + r).astype(f32))
        nc = f32(r - b)
        A = f32(nc - np.floor(nc))
        sel = np.zeros(N, bool)
        sel[run_positions(A, r, eps, qidx, Q, N)] = True
        assert not np.any(has_col & ~sel), (N, z, float(b), float(r))
        walked += int(sel.sum())
        true += int(has_col.sum())
    assert walked <= 1.5 * true + 4 * M
