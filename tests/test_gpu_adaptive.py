"""Adaptive per-pixel sample budgets (fg_render_pixelwise_adaptive_device and
fg_mlp_budget_device).  A pixel with budget b must equal, bit for bit, the
fixed render with n_samples = b and the first b offsets; the offsets of a
smaller N are a prefix of a larger N's (tests/test_sample_budget.py pins that
on the oracle), so this is also the reference's own N = b render."""
import numpy as np
import pytest

from film_grain_amd import synth
from tests import fgcases

pytestmark = pytest.mark.gpu


def _desc_with_n(desc, n):
    d = type(desc).from_buffer_copy(desc)
    d.n_samples = n
    return d


def _fixed(ctx, torch, case, n):
    offs = torch.from_numpy(np.ascontiguousarray(case.offsets_input[:n])).to("cuda:0")
    lam = torch.from_numpy(case.lam).to("cuda:0")
    out = torch.empty((case.desc.out_h, case.desc.out_w), dtype=torch.float32, device="cuda:0")
    ctx.render_device("pixel", _desc_with_n(case.desc, n), lam, offs, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _adaptive(ctx, torch, case, budget):
    offs = torch.from_numpy(case.offsets_input).to("cuda:0")
    lam = torch.from_numpy(case.lam).to("cuda:0")
    b = torch.from_numpy(budget.astype(np.uint16).view(np.int16)).to("cuda:0")
    out = torch.empty((case.desc.out_h, case.desc.out_w), dtype=torch.float32, device="cuda:0")
    ctx.render_adaptive_device(case.desc, lam, offs, b, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _check_against_fixed(ctx, torch, case, budget, name):
    got = _adaptive(ctx, torch, case, budget)
    clamped = np.clip(budget.astype(np.int64), 1, case.n_samples)
    for n in np.unique(clamped):
        want = _fixed(ctx, torch, case, int(n))
        sel = clamped == n
        bad = int(np.sum(got[sel] != want[sel]))
        print(f"[{name}] budget {n}: {int(sel.sum())} px, {bad} differ", flush=True)
        assert bad == 0, (name, int(n))


def test_uniform_budgets_equal_fixed_prefix_renders(hip_ctx, oracle):
    import torch
    case = fgcases.c2_case(oracle, w=48, h=32, n_samples=256)
    h, w = case.desc.out_h, case.desc.out_w
    for b in [1, 16, 37, 128, 255, 256]:
        _check_against_fixed(hip_ctx, torch, case, np.full((h, w), b, np.int64), f"c2 uniform {b}")
    # the oracle's own renders at N = 256 and at N = 37 (the first 37 offsets)
    got = _adaptive(hip_ctx, torch, case, np.full((h, w), 256, np.int64))
    want = oracle.render_pixelwise(case.plane, case.lam, case.offsets_input)
    assert np.array_equal(got, want)
    got = _adaptive(hip_ctx, torch, case, np.full((h, w), 37, np.int64))
    p37 = oracle.plane(case.derived, 37, case.seed, case.zoom, case.dist, case.radius)
    want = oracle.render_pixelwise(p37, case.lam, np.ascontiguousarray(case.offsets_input[:37]))
    assert np.array_equal(got, want)


def test_mixed_budgets_equal_fixed_prefix_renders(hip_ctx, oracle):
    import torch
    rng = np.random.default_rng(11)
    case = fgcases.c2_case(oracle, w=96, h=64, n_samples=256)
    h, w = case.desc.out_h, case.desc.out_w
    # per-pixel random budgets, including the clamped values 0 and > N
    budget = rng.choice(np.array([0, 1, 16, 40, 100, 256, 4000]), size=(h, w))
    _check_against_fixed(hip_ctx, torch, case, budget, "c2 random")
    # tiles at different levels: bands of budgets
    band = np.zeros((h, w), np.int64)
    band[:, : w // 3] = 16
    band[:, w // 3: 2 * w // 3] = 64
    band[:, 2 * w // 3:] = 200
    band[h // 2:, :] = np.maximum(band[h // 2:, :] // 2, 1)
    _check_against_fixed(hip_ctx, torch, case, band, "c2 bands")


def test_mixed_budgets_other_planes(hip_ctx, oracle):
    import torch
    rng = np.random.default_rng(5)
    u = synth.sinusoid_field(40, 28, 5, n_waves=3)
    cases = [
        fgcases.build_case(oracle, "const_zoom1", u, 0.12, n_samples=128, seed=3),
        fgcases.build_case(oracle, "lognorm_zoom1.5", u, 0.15, 0.15, "lognorm", zoom=1.5, n_samples=300,
                           seed=99),
        fgcases.c1_case(oracle, size=96, n_samples=64),
    ]
    for case in cases:
        h, w = case.desc.out_h, case.desc.out_w
        budget = rng.integers(1, case.n_samples + 1, size=(h, w))
        budget = np.where(rng.random((h, w)) < 0.5, budget, rng.choice([16, 32, case.n_samples], size=(h, w)))
        _check_against_fixed(hip_ctx, torch, case, budget, case.name)


def test_budget_kernel(hip_ctx):
    import torch
    from tests import mlp_ref
    rng = np.random.default_rng(2)
    n = 50000
    pred = rng.normal(-1.5, 1.0, n).astype(np.float32)
    pd = torch.from_numpy(pred).to("cuda:0")
    got = hip_ctx.mlp_budget(pd, None, n, 700.0, 16, 1024).cpu().numpy().view(np.uint16).astype(np.int64)
    want = np.clip(np.ceil(700.0 * np.exp(pred.astype(np.float64))), 16, 1024).astype(np.int64)
    # expf and the f32 product may land one step from a whole number
    assert np.all(np.abs(got - want) <= 1)
    assert np.mean(got == want) > 0.999
    assert got.min() >= 16 and got.max() <= 1024
    # the allocator without the model: s = sqrt(u (1 - u)), feature 4
    X = np.zeros((n, 16), np.float32)
    u = rng.random(n).astype(np.float32)
    X[:, 4] = u * (1 - u)
    Xb = mlp_ref.bf16(X)
    bits = (np.ascontiguousarray(Xb).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    Xd = torch.from_numpy(bits).to("cuda:0")
    got = hip_ctx.mlp_budget(None, Xd, n, 1000.0, 1, 2048).cpu().numpy().view(np.uint16).astype(np.int64)
    want = np.clip(np.ceil(1000.0 * np.sqrt(Xb[:, 4].astype(np.float64))), 1, 2048).astype(np.int64)
    assert np.all(np.abs(got - want) <= 1)
    assert np.mean(got == want) > 0.999
