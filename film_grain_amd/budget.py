"""Learned per-pixel sample budgets for the pixel-wise renderer.

The reference renders every output pixel with the same sample count N
(src/pixelwise.rs:20, `inv_samples = 1 / N`).  This module spends samples
where the Monte Carlo error is large, with the device entry points only:

  features (fg_mlp_features_device: 12 per output pixel)
  -> log std of the pixel's N = 64 error (fg_mlp_forward_device, BF16 MFMA)
  -> budgets n in [n_min, n_max] meeting a target mean squared error over the
     frame (fg_mlp_budget_device)
  -> fg_render_pixelwise_adaptive_device (pixel i is exactly the render with
     the first n_i offsets, which are the reference's own offsets for
     N = n_i: src/rng.rs:9-24 draws them as one stream).

Error model.  For a pixel whose covered fraction (over the Gaussian offset
law) is v, the estimate from the first n samples has variance s^2 / n with
s^2 = v (1 - v), and its difference to the N_REF-sample estimate (which
shares those n samples) has variance s^2 (1/n - 1/N_REF).  The network
predicts sigma_64 = s sqrt(1/64 - 1/N_REF), the spread of the N = 64 error.

Training pairs are renders the library itself makes of the same plane at
N = 64 and at N = 4096; the network is trained on the device
(fg_mlp_grad_device with the Gaussian likelihood of the residual
p64 - p4096, then fg_mlp_adam_device).  The N = 4096 render is the mean of
two 2048-sample renders over the two halves of the 4096 offsets: each half
is (covered count) * 2^-11, so their mean is (total count) * 2^-12, the
4096-sample render, exactly in f32.

Allocators compared (allocator_s, render_adaptive_frame): "learned" (the network), "binomial"
(s = sqrt(u (1 - u)) of the input intensity: no model) and "oracle" (s from
the N = 4096 render itself, which no renderer has before rendering: the
bound on what any per-pixel allocation can gain).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

N_PILOT = 64
N_REF = 4096
REF_HALF = 2048
SIGMA64_TO_S = 1.0 / math.sqrt(1.0 / N_PILOT - 1.0 / N_REF)
MODEL_PATH = Path(__file__).resolve().parent / "tuning" / "budget_mlp.npy"

# The learned-budget workload: 1920x1080 RGB planes of the sinusoid + noise +
# flat-patch field, lognormal radii mean 0.1 / std 0.05 px, filter sigma 0.8,
# zoom 1, budgets 16..1024 against fixed N = 1024.
C5 = dict(w=1920, h=1080, radius=0.1, sd=0.05, sigma=0.8, zoom=1.0, n_min=16, n_max=1024, n_fixed=1024)


@dataclass
class PlaneSpec:
    """One input plane and the parameters it is rendered with."""
    u: np.ndarray        # (in_h, in_w) f32 intensities
    channel: int         # 0..2 (the feature's one-hot)
    radius: float
    sd: float
    sigma: float
    zoom: float
    seed: int


@dataclass
class DevicePlane:
    """A plane uploaded for rendering: lambda, normalised intensities and the
    N_REF offsets (input pixels) on the device."""
    spec: PlaneSpec
    desc: object          # capi.PlaneDesc with n_samples = N_REF
    lam: object           # f32 tensor (in_h, in_w)
    u_norm: object        # f32 tensor (in_h, in_w)
    offs: object          # f32 tensor (N_REF, 2)
    out_w: int
    out_h: int


def image_planes(seed: int, w: int = C5["w"], h: int = C5["h"], radius: float = C5["radius"],
                 sd: float = C5["sd"], sigma: float = C5["sigma"], zoom: float = C5["zoom"]) -> list[PlaneSpec]:
    """The three planes of one synthetic RGB image (8 sinusoids, 2% noise,
    10% flat patches at 0.0 and 0.98 per channel)."""
    from film_grain_amd import synth
    planes = synth.c2_rgb(seed, w, h)
    return [PlaneSpec(np.ascontiguousarray(planes[c]), c, radius, sd, sigma, zoom, seed) for c in range(3)]


def mixed_planes(seed: int, w: int, h: int) -> list[PlaneSpec]:
    """Planes with drawn parameters (radius U(0.05, 0.3), sigma_r / radius
    U(0, 0.5), zoom 1 or 2): the batch workload's parameter ranges."""
    rng = np.random.default_rng(seed)
    r = float(rng.uniform(0.05, 0.3))
    sd = float(rng.uniform(0.0, 0.5)) * r
    zoom = float(rng.choice([1.0, 2.0]))
    return image_planes(seed, w, h, radius=r, sd=sd, sigma=0.8, zoom=zoom)


def upload(spec: PlaneSpec, device="cuda:0") -> DevicePlane:
    import torch
    from film_grain_amd import capi, host
    h, w = spec.u.shape
    cli = host.CliArgs(radius=spec.radius, radius_stddev=spec.sd,
                       radius_dist="lognorm" if spec.sd > 0 else "const", zoom=spec.zoom, iters=N_REF,
                       seed=spec.seed, algo="pixel", sigma=spec.sigma, color_mode="rgb")
    d, _offs, offs_in = host.derive(cli, w, h)
    desc = capi.make_desc(d.in_w, d.in_h, d.out_w, d.out_h, N_REF, d.dist_kind, d.seed, d.zoom, d.delta, d.rm,
                          d.radius_mean, d.log_mu, d.log_sigma)
    nrm, _ = host.normalize_plane(spec.u)
    lam = host.lambda_plane(nrm, d.inv_e_pi_r2)
    return DevicePlane(spec, desc, torch.from_numpy(lam).to(device),
                       torch.from_numpy(np.ascontiguousarray(nrm, dtype=np.float32)).to(device),
                       torch.from_numpy(np.ascontiguousarray(offs_in)).to(device), int(d.out_w), int(d.out_h))


def _desc_n(desc, n: int):
    d = type(desc).from_buffer_copy(desc)
    d.n_samples = n
    return d


def render_fixed(ctx, dp: DevicePlane, n: int, out=None):
    """The fixed-N render with the first n offsets (n <= 2048)."""
    import torch
    if out is None:
        out = torch.empty((dp.out_h, dp.out_w), dtype=torch.float32, device=dp.lam.device)
    ctx.render_device("pixel", _desc_n(dp.desc, n), dp.lam, dp.offs[:n], out)
    return out


def render_reference(ctx, dp: DevicePlane):
    """The N_REF = 4096 render, as the mean of the two 2048-offset halves."""
    import torch
    a = render_fixed(ctx, dp, REF_HALF)
    b = torch.empty_like(a)
    ctx.render_device("pixel", _desc_n(dp.desc, REF_HALF), dp.lam, dp.offs[REF_HALF:N_REF], b)
    return (a + b) * 0.5


def features(ctx, dp: DevicePlane, out=None):
    s = dp.spec
    return ctx.mlp_features(dp.u_norm, dp.out_w, dp.out_h, s.zoom, s.radius, s.sd, s.sigma, s.channel, out=out)


def bf16_column(X, j: int):
    """Column j of a BF16 feature tensor (int16 bits) as f32."""
    import torch
    return ((X[:, j].to(torch.int32) & 0xFFFF) << 16).view(torch.float32)


# ---------------------------------------------------------------- model ----

class ErrorModel:
    """The 12 -> 64 -> 64 -> 1 network's parameters on the device (f32, the
    layout of FG_MLP_PARAMS) with Adam moments."""

    def __init__(self, params=None, device="cuda:0", seed: int = 0):
        import torch
        from film_grain_amd import capi
        if params is None:
            params = init_params(seed)
        self.P = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).to(device)
        assert self.P.numel() == capi.MLP_PARAMS
        self.m = torch.zeros_like(self.P)
        self.v = torch.zeros_like(self.P)
        self.step = 0

    def predict(self, ctx, X, out=None):
        """Predicted log sigma_64 per row."""
        return ctx.mlp_forward(self.P, X, out=out)

    def numpy(self) -> np.ndarray:
        return self.P.detach().cpu().numpy().copy()

    def save(self, path: Path = MODEL_PATH) -> None:
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        np.save(path, self.numpy())

    @classmethod
    def load(cls, path: Path = MODEL_PATH, device="cuda:0") -> "ErrorModel":
        return cls(np.load(path), device=device)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, padding inputs 12..15 zero; the output
    bias starts at the log sigma_64 of a mid-grey pixel."""
    from film_grain_amd import capi
    rng = np.random.default_rng(seed)
    P = np.zeros(capi.MLP_PARAMS, np.float32)
    W1 = rng.normal(0, math.sqrt(2 / 12), (16, 64)).astype(np.float32)
    W1[12:] = 0.0
    P[0:1024] = W1.ravel()
    P[1088:5184] = rng.normal(0, math.sqrt(2 / 64), 64 * 64)
    P[5248:5312] = rng.normal(0, math.sqrt(1 / 64), 64)
    P[5312] = math.log(0.5 / SIGMA64_TO_S)
    return P


def build_training_set(ctx, specs, rows_per_plane: int, seed: int = 0, log=None):
    """Features and N = 64 residuals (p64 - p4096) of rows_per_plane random
    output pixels per plane.  Returns (X int16 [R, 16], e f32 [R])."""
    import torch
    g = torch.Generator(device="cuda:0")
    g.manual_seed(seed)
    Xs, es = [], []
    t0 = time.time()
    for i, spec in enumerate(specs):
        dp = upload(spec)
        X = features(ctx, dp)
        p64 = render_fixed(ctx, dp, N_PILOT)
        ref = render_reference(ctx, dp)
        e = (p64 - ref).reshape(-1)
        npx = e.numel()
        idx = torch.randint(0, npx, (min(rows_per_plane, npx),), device="cuda:0", generator=g)
        Xs.append(X[idx])
        es.append(e[idx])
        if log and (i % 30 == 0 or i == len(specs) - 1):
            log(f"[budget] training planes {i + 1}/{len(specs)} ({time.time() - t0:.1f} s)")
    return torch.cat(Xs), torch.cat(es)


def train(ctx, model: ErrorModel, X, e, steps: int = 3000, batch: int = 65536, lr: float = 3e-3,
          seed: int = 0, log=None):
    """Adam on the Gaussian likelihood of the residuals e (minibatches drawn
    uniformly).  Returns the mean loss of the first and last 50 steps."""
    import torch
    from film_grain_amd import capi
    g = torch.Generator(device="cuda:0")
    g.manual_seed(seed)
    grad = torch.zeros(capi.MLP_PARAMS, dtype=torch.float32, device="cuda:0")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    losses = torch.zeros(steps, dtype=torch.float32, device="cuda:0")
    R = X.shape[0]
    for k in range(steps):
        idx = torch.randint(0, R, (batch,), device="cuda:0", generator=g)
        xb, eb = X[idx], e[idx]
        grad.zero_()
        loss.zero_()
        ctx.mlp_grad(model.P, xb, eb, grad, loss, loss_kind="nll")
        model.step += 1
        # step decay: full rate for 60%, then 1/3, then 1/10
        rate = lr if k < 0.6 * steps else (lr / 3 if k < 0.85 * steps else lr / 10)
        ctx.mlp_adam(model.P, grad, model.m, model.v, rate, model.step)
        losses[k] = loss[0] / batch
        if log and (k % 1000 == 0 or k == steps - 1):
            log(f"[budget] step {k}: nll {float(losses[k]):.4f}")
    h = losses.cpu().numpy()
    w = min(50, steps)
    return float(h[:w].mean()), float(h[-w:].mean())


def nll(ctx, model: ErrorModel, X, e) -> float:
    """Mean Gaussian negative log-likelihood of residuals e."""
    import torch
    y = model.predict(ctx, X)
    return float(torch.mean(y + 0.5 * e * e * torch.exp(torch.clamp(-2 * y, -80, 80))))


# ------------------------------------------------------------ allocation ----

def predicted_mse(s, n):
    """Expected squared difference to the N_REF render of per-sample stds s
    rendered with budgets n."""
    return s * s * (1.0 / n - 1.0 / N_REF)


def choose_scale(s, target_mse: float, n_min: int, n_max: int, probes: int = 64, rounds: int = 3,
                 max_rows: int = 1 << 18) -> float:
    """The smallest c whose budgets clamp(ceil(c s), n_min, n_max) have mean
    predicted_mse <= target_mse (on a strided subsample of at most max_rows
    rows), by rounds of `probes` geometric candidates.  Returns c (a target
    below what n_max everywhere reaches gives the c that puts n_max
    everywhere)."""
    import torch
    s = s.reshape(-1).float()
    step = max(1, s.numel() // max_rows)
    ss = s[::step]
    lo, hi = 1e-2, 2.0 * n_max / max(float(ss[ss > 0].min()) if bool((ss > 0).any()) else 1.0, 1e-6)
    for _ in range(rounds):
        cs = torch.exp(torch.linspace(math.log(lo), math.log(hi), probes, device=ss.device))
        n = torch.clamp(torch.ceil(cs[:, None] * ss[None, :]), n_min, n_max)
        m = predicted_mse(ss[None, :], n).mean(dim=1)
        ok = (m <= target_mse).nonzero()
        if ok.numel() == 0:
            return float(cs[-1])
        k = int(ok[0])
        if k == 0:
            return float(cs[0])
        lo, hi = float(cs[k - 1]), float(cs[k])
    return hi


def allocator_s(kind: str, model: ErrorModel | None, ctx, X, ref=None):
    """Per-sample std estimates of one plane for an allocator kind."""
    import torch
    if kind == "learned":
        return torch.exp(model.predict(ctx, X)) * SIGMA64_TO_S
    if kind == "binomial":
        return torch.sqrt(torch.clamp(bf16_column(X, 4), min=0.0))
    if kind == "oracle":
        r = ref.reshape(-1)
        return torch.sqrt(torch.clamp(r * (1 - r), min=0.0))
    raise ValueError(kind)


@dataclass
class FrameBuffers:
    """Per-plane device buffers of the adaptive renderer, allocated once."""
    X: list
    pred: list
    budget: list
    out: list


def frame_buffers(planes: list[DevicePlane]) -> FrameBuffers:
    import torch
    from film_grain_amd import capi
    dev = planes[0].lam.device
    X, pred, budget, out = [], [], [], []
    # the planes' predictions are views of one array: the scale search reads
    # the whole frame in one call
    pred_all = torch.empty(sum(dp.out_w * dp.out_h for dp in planes), dtype=torch.float32, device=dev)
    at = 0
    for dp in planes:
        n = dp.out_w * dp.out_h
        X.append(torch.empty((n, capi.MLP_FEATURES), dtype=torch.int16, device=dev))
        pred.append(pred_all[at:at + n])
        at += n
        budget.append(torch.empty(n, dtype=torch.int16, device=dev))
        out.append(torch.empty((dp.out_h, dp.out_w), dtype=torch.float32, device=dev))
    return FrameBuffers(X, pred, budget, out)


def scale_for_mean_budget(s, mean_budget: float, n_min: int, n_max: int, probes: int = 64, rounds: int = 3,
                          max_rows: int = 1 << 18) -> float:
    """The largest c whose budgets clamp(ceil(c s), n_min, n_max) average at
    most mean_budget (strided subsample of at most max_rows rows): the
    allocation of a fixed total of samples."""
    import torch
    s = s.reshape(-1).float()
    ss = s[::max(1, s.numel() // max_rows)]
    pos = ss[ss > 0]
    lo = 1e-3
    hi = 2.0 * n_max / max(float(pos.min()), 1e-6) if pos.numel() else 1.0
    for _ in range(rounds):
        cs = torch.exp(torch.linspace(math.log(lo), math.log(hi), probes, device=ss.device))
        m = torch.clamp(torch.ceil(cs[:, None] * ss[None, :]), n_min, n_max).mean(dim=1)
        ok = (m <= mean_budget).nonzero()
        if ok.numel() == 0:
            return float(cs[0])
        k = int(ok[-1])
        if k == probes - 1:
            return float(cs[-1])
        lo, hi = float(cs[k]), float(cs[k + 1])
    return lo


def render_adaptive_frame(ctx, model: ErrorModel | None, planes: list[DevicePlane], bufs: FrameBuffers,
                          target_mse: float | None = None, kind: str = "learned", n_min: int = C5["n_min"],
                          n_max: int = C5["n_max"], refs=None, mean_budget: float | None = None):
    """Planes rendered with per-pixel budgets under one scale c for all of
    them: features, predictions, c (meeting target_mse, or averaging
    mean_budget samples per pixel), budgets, adaptive renders, each launched
    for the largest budget of its plane.  Returns (outputs, c)."""
    import torch
    if kind == "learned" and target_mse is not None and mean_budget is None:
        return _render_learned_frame(ctx, model, planes, bufs, target_mse, n_min, n_max)
    s_all = []
    for i, dp in enumerate(planes):
        features(ctx, dp, out=bufs.X[i])
        if kind == "learned":
            model.predict(ctx, bufs.X[i], out=bufs.pred[i])
            s_all.append(torch.exp(bufs.pred[i]) * SIGMA64_TO_S)
        else:
            s_all.append(allocator_s(kind, model, ctx, bufs.X[i], None if refs is None else refs[i]))
    s_cat = torch.cat(s_all)
    if mean_budget is not None:
        c = scale_for_mean_budget(s_cat, mean_budget, n_min, n_max)
    else:
        c = choose_scale(s_cat, target_mse, n_min, n_max)
    for i, dp in enumerate(planes):
        n = dp.out_w * dp.out_h
        if kind == "learned":
            ctx.mlp_budget(bufs.pred[i], None, n, c * SIGMA64_TO_S, n_min, n_max, out=bufs.budget[i])
        elif kind == "binomial":
            ctx.mlp_budget(None, bufs.X[i], n, c, n_min, n_max, out=bufs.budget[i])
        else:
            ctx.mlp_budget(torch.log(torch.clamp(s_all[i], min=1e-30)).contiguous(), None, n, c, n_min, n_max,
                           out=bufs.budget[i])
    # the largest budget of a plane sizes its launch (tile shape, sample levels)
    his = torch.stack([(b.to(torch.int32) & 0xFFFF).max() for b in bufs.budget]).tolist()
    for i, dp in enumerate(planes):
        n_hi = max(1, min(n_max, int(his[i])))
        ctx.render_adaptive_device(_desc_n(dp.desc, n_hi), dp.lam, dp.offs[:n_hi], bufs.budget[i], bufs.out[i])
    return bufs.out, c


def _render_learned_frame(ctx, model: ErrorModel, planes: list[DevicePlane], bufs: FrameBuffers, target_mse: float,
                          n_min: int, n_max: int):
    """render_adaptive_frame for the learned model at a target error, with no
    host round trip: the scale search (choose_scale's rounds, on the device)
    and the budgets read the scale from device memory, and every plane's
    launch is sized for n_max (its tiles pick their own levels).  Returns
    (outputs, c) with c a one-element device tensor."""
    for i, dp in enumerate(planes):
        features(ctx, dp, out=bufs.X[i])
        model.predict(ctx, bufs.X[i], out=bufs.pred[i])
    import torch
    base = bufs.pred[0]._base
    if base is not None and base.numel() == sum(p.numel() for p in bufs.pred):
        pred_all = base
    else:
        pred_all = torch.cat(bufs.pred)
    sc = ctx.budget_scale(pred_all, SIGMA64_TO_S, target_mse, n_min, n_max, N_REF)
    for i, dp in enumerate(planes):
        ctx.mlp_budget_scale(bufs.pred[i], None, dp.out_w * dp.out_h, sc[1:], n_min, n_max, out=bufs.budget[i])
    for i, dp in enumerate(planes):
        ctx.render_adaptive_device(_desc_n(dp.desc, n_max), dp.lam, dp.offs[:n_max], bufs.budget[i], bufs.out[i])
    return bufs.out, sc[:1]


# ------------------------------------------------- tile-level allocation ----
#
# The adaptive renderer (fg_render_pixelwise_adaptive_device) renders each
# tile at one sample level: the first of n_min, 2 n_min, ..., n_max that
# covers the largest budget among the tile's pixels, and its work is that
# level times the tile's per-sample cost.  A pixel budget below its tile's
# level saves nothing, so the allocation that matches the kernel gives every
# pixel of tile t the level L_t itself, and picks the levels that minimise the
# frame's work  sum_t c_t L_t  subject to its predicted error
#     sum_t S_t (1 / L_t - 1 / N_REF) <= target * pixels,
# with S_t the sum of s^2 over the tile and c_t its per-sample cost.  The
# Lagrange condition gives L_t = sqrt(mu S_t / c_t), the n ~ s / sqrt(c) rule
# at the kernel's granularity; with discrete levels each tile takes the level
# minimising c_t L + mu S_t / L, and mu is the smallest meeting the target.
# The per-sample cost of a pixel is a + b lambda (the splat's work per sample
# has a fixed part and a part that grows with the grains its window holds),
# c_t = sum over the tile of 1 + kappa lambda, kappa = b / a measured by
# calibrate_cost.


def tile_levels_list(n_min: int, n_max: int) -> list[int]:
    """The adaptive renderer's sample levels: n_min doubling below n_max,
    then n_max."""
    out, level = [], n_min
    while level < n_max:
        out.append(level)
        level *= 2
    return out + [n_max]


def out_lambda(dp: DevicePlane):
    """lambda at each output pixel's input position (nearest pixel)."""
    import torch
    z = dp.spec.zoom
    dev = dp.lam.device
    ys = torch.clamp(torch.floor((torch.arange(dp.out_h, device=dev) + 0.5) / z).long(), 0, dp.lam.shape[0] - 1)
    xs = torch.clamp(torch.floor((torch.arange(dp.out_w, device=dev) + 0.5) / z).long(), 0, dp.lam.shape[1] - 1)
    return dp.lam[ys][:, xs]


def tile_sums(x, tw: int, th: int):
    """Sums of an (h, w) tensor over tw x th tiles (edge tiles partial)."""
    import torch
    h, w = x.shape
    ny, nx = (h + th - 1) // th, (w + tw - 1) // tw
    p = torch.zeros((ny * th, nx * tw), dtype=x.dtype, device=x.device)
    p[:h, :w] = x
    return p.view(ny, th, nx, tw).sum(dim=(1, 3))


def calibrate_cost(ctx, spec: PlaneSpec | None = None, us=(0.3, 0.9), ns=(64, 512), reps: int = 3) -> dict:
    """kappa = b / a of the per-sample cost a + b lambda of a pixel, from the
    fixed-N render time of two planes of uniform intensity (one pixel at 1 so
    that normalisation keeps the level) at two sample counts: the slope in N
    removes the per-plane generation, and the two lambdas separate a and b."""
    spec = spec or image_planes(1)[0]
    slopes, lams = [], []
    for u in us:
        plane = np.full_like(spec.u, u)
        plane[0, 0] = 1.0
        dp = upload(PlaneSpec(plane, spec.channel, spec.radius, spec.sd, spec.sigma, spec.zoom, spec.seed))
        out = None
        ts = []
        for n in ns:
            def run():
                render_fixed(ctx, dp, n, out=out)
            out = render_fixed(ctx, dp, n)
            ts.append(_time_ms(run, reps))
        slopes.append((ts[1] - ts[0]) / (ns[1] - ns[0]))
        lams.append(float(dp.lam[dp.lam.shape[0] // 2, dp.lam.shape[1] // 2]))
        del dp
    b = (slopes[1] - slopes[0]) / (lams[1] - lams[0])
    a = slopes[0] - b * lams[0]
    kappa = b / a if a > 0 and b > 0 else 0.0
    return dict(kappa=kappa, ms_per_sample=slopes, lambdas=lams)


def choose_tile_levels(S, C, pixels: int, target_mse: float, levels, n_ref: int = N_REF, probes: int = 64,
                       rounds: int = 3):
    """Per tile the level minimising C L + mu S / L for the smallest mu whose
    predicted frame error sum S (1/L - 1/n_ref) / pixels is <= target_mse
    (rounds of `probes` geometric candidates).  S, C: 1-D tile tensors.
    Returns (levels per tile, mu); a target below what the top level reaches
    everywhere gives the top level everywhere."""
    import torch
    L = torch.tensor(levels, dtype=torch.float64, device=S.device)
    inv = 1.0 / L
    S, C = S.double(), C.double()

    def pick(mu):  # mu: (k,) -> level index (k, T), frame mse (k,)
        j = torch.argmin(C[None, :, None] * L[None, None, :] + mu[:, None, None] * S[None, :, None] * inv, dim=2)
        mse = (S[None, :] * (inv[j] - 1.0 / n_ref)).sum(dim=1) / pixels
        return j, mse

    pos = S > 0
    if not bool(pos.any()):
        return torch.full_like(S, levels[0]), 0.0
    ratio = C[pos] / S[pos]
    lo = float(ratio.min()) * levels[0] ** 2 * 0.25
    hi = float(ratio.max()) * levels[-1] ** 2 * 4.0
    for _ in range(rounds):
        mus = torch.exp(torch.linspace(math.log(lo), math.log(hi), probes, device=S.device, dtype=torch.float64))
        _, mse = pick(mus)
        ok = (mse <= target_mse).nonzero()
        if ok.numel() == 0:
            lo = hi = float(mus[-1])
            break
        k = int(ok[0])
        if k == 0:
            lo = hi = float(mus[0])
            break
        lo, hi = float(mus[k - 1]), float(mus[k])
    j, _ = pick(torch.tensor([hi], dtype=torch.float64, device=S.device))
    return L[j[0]], hi


def render_tiled_frame(ctx, model: ErrorModel | None, planes: list[DevicePlane], bufs: FrameBuffers,
                       target_mse: float, kappa: float, kind: str = "learned", n_min: int = C5["n_min"],
                       n_max: int = C5["n_max"], refs=None):
    """The frame with tile-level budgets (the allocation above): features,
    per-pixel spreads, per-tile S and c, the levels, budgets equal to the
    tile's level, and the adaptive renders.  The launch is sized for the
    largest level the frame uses (smaller per-sample bitsets allow larger
    tiles): the levels are chosen on the tiles of an n_max launch, and again
    on the tiles of a launch for the top level they use, when those differ.
    Returns (outputs, info)."""
    import torch
    s_list = []
    for i, dp in enumerate(planes):
        features(ctx, dp, out=bufs.X[i])
        if kind == "learned":
            model.predict(ctx, bufs.X[i], out=bufs.pred[i])
            s = torch.exp(bufs.pred[i]) * SIGMA64_TO_S
        else:
            s = allocator_s(kind, model, ctx, bufs.X[i], None if refs is None else refs[i])
        s_list.append((s.reshape(dp.out_h, dp.out_w).float()) ** 2)
    cost_px = [1.0 + kappa * out_lambda(dp) for dp in planes]
    pixels = sum(dp.out_w * dp.out_h for dp in planes)

    def allocate(n_top):
        levels = tile_levels_list(n_min, n_top)
        S_all, C_all, shapes = [], [], []
        for i, dp in enumerate(planes):
            cfg = ctx.launch_config(_desc_n(dp.desc, n_top), "pixel")
            tw, th = cfg["pix_tile_w"], cfg["pix_tile_h"]
            S = tile_sums(s_list[i], tw, th)
            S_all.append(S.reshape(-1))
            C_all.append(tile_sums(cost_px[i], tw, th).reshape(-1))
            shapes.append((S.shape, tw, th))
        lv, mu = choose_tile_levels(torch.cat(S_all), torch.cat(C_all), pixels, target_mse, levels)
        return lv, mu, shapes, torch.cat(C_all)

    n_top = n_max
    lv, mu, shapes, C = allocate(n_top)
    top = int(lv.max())
    if top < n_max:
        shapes_top = [ctx.launch_config(_desc_n(dp.desc, top), "pixel") for dp in planes]
        if any((c["pix_tile_w"], c["pix_tile_h"]) != (sh[1], sh[2]) for c, sh in zip(shapes_top, shapes)):
            n_top = top
            lv, mu, shapes, C = allocate(n_top)
        else:
            n_top = top
    at = 0
    for i, dp in enumerate(planes):
        (ny, nx), tw, th = shapes[i]
        t = lv[at:at + ny * nx].reshape(ny, nx).to(torch.int16)
        at += ny * nx
        full = t.repeat_interleave(th, dim=0).repeat_interleave(tw, dim=1)[:dp.out_h, :dp.out_w]
        bufs.budget[i].copy_(full.reshape(-1))
    for i, dp in enumerate(planes):
        ctx.render_adaptive_device(_desc_n(dp.desc, n_top), dp.lam, dp.offs[:n_top], bufs.budget[i], bufs.out[i])
    return bufs.out, dict(mu=mu, work=float((C.double() * lv).sum()), n_top=n_top)


def _frame_rmse(outs, refs) -> float:
    import torch
    se = sum(float(torch.sum((o - r).double() ** 2)) for o, r in zip(outs, refs))
    n = sum(r.numel() for r in refs)
    return math.sqrt(se / n)


def _time_ms(fn, reps: int = 3) -> float:
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts))


def fixed_curve(ctx, planes, refs, ns=(64, 128, 256, 512, 1024, 2048), reps: int = 3):
    """(N, ms/frame, rmse vs N_REF) of fixed-N rendering."""
    import torch
    outs = [torch.empty((dp.out_h, dp.out_w), dtype=torch.float32, device=dp.lam.device) for dp in planes]
    rows = []
    for n in ns:
        def run():
            for dp, o in zip(planes, outs):
                render_fixed(ctx, dp, n, out=o)
        ms = _time_ms(run, reps)
        rows.append((n, ms, _frame_rmse(outs, refs)))
    return rows


def ms_at_rmse(curve, rmse: float) -> float:
    """Fixed-N ms/frame at a given RMSE, interpolated log-log between the
    measured sample counts (None outside the measured range)."""
    pts = sorted(curve, key=lambda r: -r[2])  # decreasing rmse
    for (n0, t0, e0), (n1, t1, e1) in zip(pts, pts[1:]):
        if e1 <= rmse <= e0:
            a = (math.log(rmse) - math.log(e0)) / (math.log(e1) - math.log(e0)) if e0 != e1 else 0.0
            return math.exp(math.log(t0) + a * (math.log(t1) - math.log(t0)))
    return None


def evaluate(ctx, model: ErrorModel | None, specs_per_frame, targets_n=(256, 1024),
             kinds=("learned", "binomial", "oracle"), reps: int = 3, log=None):
    """Per frame: the fixed-N curve and, for each allocator and each target
    (the RMSE fixed N reaches), the adaptive frame's ms, RMSE and mean
    budget, with the fixed-N ms at the same RMSE."""
    import torch
    results = []
    for fi, specs in enumerate(specs_per_frame):
        planes = [upload(s) for s in specs]
        refs = [render_reference(ctx, dp) for dp in planes]
        curve = fixed_curve(ctx, planes, refs, reps=reps)
        bufs = frame_buffers(planes)
        frame = {"fixed": [dict(n=n, ms=ms, rmse=e) for n, ms, e in curve], "adaptive": []}
        for tn in targets_n:
            target_rmse = next(e for n, ms, e in curve if n == tn)
            for kind in kinds:
                if kind == "learned" and model is None:
                    continue

                def run():
                    render_adaptive_frame(ctx, model, planes, bufs, target_rmse ** 2, kind, refs=refs)
                ms = _time_ms(run, reps)
                outs, c = render_adaptive_frame(ctx, model, planes, bufs, target_rmse ** 2, kind, refs=refs)
                torch.cuda.synchronize()
                rmse = _frame_rmse(outs, refs)
                mean_n = float(np.mean([float((b.to(torch.int32) & 0xFFFF).float().mean())
                                        for b in bufs.budget]))
                fixed_ms = ms_at_rmse(curve, rmse)
                row = dict(target_n=tn, kind=kind, ms=ms, rmse=rmse, mean_budget=mean_n, c=float(c),
                           fixed_ms_at_rmse=fixed_ms, speedup=(fixed_ms / ms) if fixed_ms else None)
                frame["adaptive"].append(row)
                if log:
                    log(f"[budget] frame {fi} target N={tn} {kind}: {ms:.2f} ms rmse {rmse * 255:.3f}/255 "
                        f"mean n {mean_n:.1f} fixed at same rmse {fixed_ms if fixed_ms is None else round(fixed_ms, 2)}"
                        f" ms")
        results.append(frame)
        del planes, refs, bufs
    return results


def compare_at_equal_budget(ctx, model: ErrorModel | None, specs, mean_budgets=(64, 256),
                            kinds=("learned", "binomial", "oracle"), n_min: int = C5["n_min"],
                            n_max: int = C5["n_max"], log=None):
    """All planes of `specs` rendered together under one sample total (the
    mean budget per pixel): per allocator, the RMSE against N_REF and the
    mean budget reached.  A better error model puts the same samples where
    they reduce the error most."""
    planes = [upload(s) for s in specs]
    refs = [render_reference(ctx, dp) for dp in planes]
    bufs = frame_buffers(planes)
    import torch
    rows = []
    for mb in mean_budgets:
        for kind in kinds:
            if kind == "learned" and model is None:
                continue
            outs, c = render_adaptive_frame(ctx, model, planes, bufs, kind=kind, n_min=n_min, n_max=n_max,
                                            refs=refs, mean_budget=mb)
            torch.cuda.synchronize()
            mean_n = float(np.mean([float((b.to(torch.int32) & 0xFFFF).float().mean()) for b in bufs.budget]))
            row = dict(mean_budget=mb, kind=kind, rmse=_frame_rmse(outs, refs), mean_n=mean_n)
            rows.append(row)
            if log:
                log(f"[budget] equal budget {mb}: {kind} rmse {row['rmse'] * 255:.3f}/255 mean n {mean_n:.1f}")
    fixed = []
    for mb in mean_budgets:
        outs = [render_fixed(ctx, dp, int(mb)) for dp in planes]
        fixed.append(dict(mean_budget=mb, kind="fixed", rmse=_frame_rmse(outs, refs), mean_n=float(mb)))
    return rows + fixed


def s_ratio(refs) -> float:
    """E[s^2] / E[s]^2 of the per-sample stds of N_REF renders: the largest
    factor by which any per-pixel allocation can cut the mean sample count
    at equal predicted error (budgets unclamped, N_REF terms neglected)."""
    import torch
    s = torch.cat([torch.sqrt(torch.clamp(r * (1 - r), min=0)).reshape(-1) for r in refs]).double()
    return float((s * s).mean() / s.mean() ** 2)


def main(argv=None):
    """Trains the error model and evaluates the learned budget:
        python -m film_grain_amd.budget train [--images 200] [--save]
        python -m film_grain_amd.budget eval [--frames 2]
    """
    import argparse
    import json
    import sys

    import torch

    from film_grain_amd import capi
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--images", type=int, default=200)
    ap.add_argument("--rows", type=int, default=65536, help="training pixels per plane")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--frames", type=int, default=2, help="held-out frames to evaluate")
    ap.add_argument("--width", type=int, default=C5["w"])
    ap.add_argument("--height", type=int, default=C5["h"])
    ap.add_argument("--mixed", action="store_true", help="draw radius / sigma_r / zoom per image")
    ap.add_argument("--model", default=str(MODEL_PATH))
    ap.add_argument("--save", action="store_true")
    ap.add_argument("--json", default=None, help="write the evaluation here")
    ap.add_argument("--equal-budget", action="store_true",
                    help="compare allocators at equal total samples over all frames together")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    ctx = capi.Context(0)
    make = (lambda s: mixed_planes(s, args.width, args.height)) if args.mixed else \
        (lambda s: image_planes(s, args.width, args.height))
    if args.cmd == "train":
        specs = [p for i in range(args.images) for p in make(10_000 + i)]
        X, e = build_training_set(ctx, specs, args.rows, log=log)
        held = [p for i in range(4) for p in make(90_000 + i)]
        Xh, eh = build_training_set(ctx, held, args.rows, seed=1)
        model = ErrorModel()
        before = nll(ctx, model, Xh, eh)
        first, last = train(ctx, model, X, e, steps=args.steps, log=log)
        after = nll(ctx, model, Xh, eh)
        # the likelihood of the binomial model with its best single scale
        # (floored at s = 0.02, the spread of a nearly flat pixel)
        sb = torch.clamp(torch.sqrt(torch.clamp(bf16_column(Xh, 4), min=0.0)), min=0.02)
        k2 = float(torch.mean(eh * eh / (sb * sb)))
        yb = torch.log(sb * math.sqrt(k2))
        nll_binomial = float(torch.mean(yb + 0.5 * eh * eh * torch.exp(-2 * yb)))
        log(f"[budget] held-out nll: init {before:.4f} trained {after:.4f} binomial(best scale) "
            f"{nll_binomial:.4f}; train {first:.4f} -> {last:.4f}")
        if args.save:
            model.save(Path(args.model))
        print(json.dumps(dict(rows=int(X.shape[0]), held_out_nll=after, held_out_nll_binomial=nll_binomial,
                              train_nll_first=first, train_nll_last=last)))
    else:
        model = ErrorModel.load(Path(args.model)) if Path(args.model).exists() else None
        frames = [make(50_000 + i) for i in range(args.frames)]
        if args.equal_budget:
            res = compare_at_equal_budget(ctx, model, [p for f in frames for p in f], log=log)
        else:
            res = evaluate(ctx, model, frames, log=log)
        out = json.dumps(res, indent=1)
        if args.json:
            Path(args.json).write_text(out + "\n")
        print(out)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
