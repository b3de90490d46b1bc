"""Per-pixel error model on the GPU (fg_mlp_*): features, the BF16 MFMA
forward pass and the fused backward pass against the CPU fp32 reference
(tests/mlp_ref.py), and a short training run."""
import numpy as np
import pytest

from film_grain_amd import capi, synth
from tests import mlp_ref

pytestmark = pytest.mark.gpu


def _params(rng, scale=0.3):
    P = rng.normal(0, scale, mlp_ref.N_PARAMS).astype(np.float32)
    P[mlp_ref.P_W1 + 12 * 64:mlp_ref.P_B1] = 0.0
    return P


def _features(rng, n):
    X = np.zeros((n, 16), np.float32)
    X[:, :12] = rng.uniform(-1, 1, (n, 12))
    return mlp_ref.bf16(X)


def _to_bf16_tensor(torch, X):
    bits = np.ascontiguousarray(mlp_ref.bf16(X)).view(np.uint32) >> 16
    return torch.from_numpy(bits.astype(np.uint16).view(np.int16)).to("cuda:0")


@pytest.mark.parametrize("n,kind", [(16, "squared"), (1000, "squared"), (40000, "squared"), (1000, "nll"),
                                    (40000, "nll")])
def test_forward_and_gradient_match_cpu_fp32(hip_ctx, n, kind):
    import torch
    rng = np.random.default_rng(n)
    P = _params(rng)
    X = _features(rng, n)
    t = rng.normal(0, 1, n).astype(np.float32)
    Pd = torch.from_numpy(P).to("cuda:0")
    Xd = _to_bf16_tensor(torch, X)
    td = torch.from_numpy(t).to("cuda:0")
    y = hip_ctx.mlp_forward(Pd, Xd).cpu().numpy()
    y_ref = mlp_ref.forward(P, X)[-1]
    err = np.abs(y - y_ref).max() / (np.abs(y_ref).max() + 1e-6)
    print(f"[mlp n={n}] forward max rel err {err:.3g}", flush=True)
    assert err < 1e-3
    grad = torch.zeros(mlp_ref.N_PARAMS, dtype=torch.float32, device="cuda:0")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    hip_ctx.mlp_grad(Pd, Xd, td, grad, loss, loss_kind=kind)
    g = grad.cpu().numpy()
    g_ref, l_ref = mlp_ref.grad(P, X, t, loss=kind)
    for name, a, b in [("W1", mlp_ref.P_W1, mlp_ref.P_B1), ("b1", mlp_ref.P_B1, mlp_ref.P_W2),
                       ("W2", mlp_ref.P_W2, mlp_ref.P_B2), ("b2", mlp_ref.P_B2, mlp_ref.P_W3),
                       ("w3", mlp_ref.P_W3, mlp_ref.P_B3), ("b3", mlp_ref.P_B3, mlp_ref.N_PARAMS)]:
        scale = np.abs(g_ref[a:b]).max() + 1e-12
        e = np.abs(g[a:b] - g_ref[a:b]).max() / scale
        print(f"[mlp n={n} {kind}] grad {name} max rel err {e:.3g}", flush=True)
        assert e < 5e-3, name
    assert abs(float(loss.item()) - l_ref) <= 1e-4 * abs(l_ref) + 1e-6


def test_features_match_cpu(hip_ctx):
    import torch
    u = synth.sinusoid_field(37, 23, 4, n_waves=3)
    for zoom, ch in [(1.0, 0), (2.0, 1), (0.5, 2), (1.7, 0)]:
        out_w, out_h = max(1, round(37 * zoom)), max(1, round(23 * zoom))
        Xd = hip_ctx.mlp_features(torch.from_numpy(u).to("cuda:0"), out_w, out_h, zoom, 0.1, 0.05, 0.8, ch)
        bits = Xd.cpu().numpy().view(np.uint16)
        got = mlp_ref.bf16_bits_to_f32(bits)
        want = mlp_ref.bf16(mlp_ref.features(u, out_w, out_h, zoom, 0.1, 0.05, 0.8, ch))
        # float32 sums in a different order: one BF16 step, plus the
        # cancellation of the variance near zero (its square root)
        d = np.abs(got - want)
        assert np.all(d <= np.abs(want) * 2.0 ** -7 + 1e-3), (zoom, ch, d.max())


def test_training_reduces_loss(hip_ctx):
    """Adam on the device fits a smooth target of the features."""
    import torch
    rng = np.random.default_rng(3)
    n = 65536
    X = _features(rng, n)
    t = (np.log(np.sqrt(np.clip(0.25 - 0.2 * X[:, 0] ** 2, 1e-4, None))) + 0.3 * X[:, 3]).astype(np.float32)
    P = _params(rng, 0.1)
    dev = "cuda:0"
    Pd, Xd, td = torch.from_numpy(P).to(dev), _to_bf16_tensor(torch, X), torch.from_numpy(t).to(dev)
    m = torch.zeros_like(Pd)
    v = torch.zeros_like(Pd)
    grad = torch.zeros_like(Pd)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    losses = []
    for step in range(1, 301):
        grad.zero_()
        loss.zero_()
        hip_ctx.mlp_grad(Pd, Xd, td, grad, loss)
        hip_ctx.mlp_adam(Pd, grad, m, v, 3e-3, step)
        if step in (1, 300):
            losses.append(float(loss.item()) / n)
    print(f"[mlp training] loss {losses[0]:.4g} -> {losses[1]:.4g}", flush=True)
    assert losses[1] < 0.2 * losses[0]
    assert float(Pd[mlp_ref.P_W1 + 12 * 64:mlp_ref.P_B1].abs().max()) == 0.0


def test_nll_training_recovers_the_noise_scale(hip_ctx):
    """Trained on single residuals e ~ N(0, s(x)^2), the network's output
    approaches log s(x): the Gaussian likelihood fits the error scale from
    one draw per row, as the budget trainer uses it."""
    import torch
    rng = np.random.default_rng(4)
    n = 131072
    X = _features(rng, n)
    log_s = (-3.0 + 1.2 * X[:, 0] - 0.8 * np.abs(X[:, 1])).astype(np.float64)
    e = (np.exp(log_s) * rng.standard_normal(n)).astype(np.float32)
    P = _params(rng, 0.1)
    dev = "cuda:0"
    Pd, Xd, ed = torch.from_numpy(P).to(dev), _to_bf16_tensor(torch, X), torch.from_numpy(e).to(dev)
    m, v, grad = torch.zeros_like(Pd), torch.zeros_like(Pd), torch.zeros_like(Pd)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    for step in range(1, 601):
        grad.zero_()
        loss.zero_()
        hip_ctx.mlp_grad(Pd, Xd, ed, grad, loss, loss_kind="nll")
        hip_ctx.mlp_adam(Pd, grad, m, v, 1e-2 if step < 400 else 3e-3, step)
    y = hip_ctx.mlp_forward(Pd, Xd).cpu().numpy().astype(np.float64)
    err = np.abs(y - log_s)
    print(f"[mlp nll] mean |log s err| {err.mean():.3f}, loss {float(loss.item()) / n:.4f}", flush=True)
    assert err.mean() < 0.1
