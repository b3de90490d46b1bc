"""The device's f64 math against the C library, on MI355X.

The device evaluates the reference's exponentials with fg_exp_dd (fg_exp.h),
which equals glibc's exp except on about 2 inputs in 10^4, and corrects the
Knuth threshold exp(-x) on its whole f32 domain with the build-time table of
those inputs (fg_exp_glibc.h); log, tan and sqrt are the device library's.

The reference calls glibc (through Rust's f64::exp/ln/tan/sqrt and
rand_distr) for the Knuth threshold exp(-lambda) (Poisson::new,
src/pixelwise.rs:81, src/grainwise.rs:42), the LogNormal radius
exp(mu + sigma z) cast to f32 (src/model.rs:137-148), and the rejection
sampler's tan, log, exp and log_gamma (lambda >= 12: C3's pixels).  The device
evaluates the same expressions (fg_math_device runs them as the kernels do);
the oracle calls glibc (fgo_glibc_math).

* exp(-x) over every f32 x in (0, 12): the whole Knuth-threshold domain.
* The f32 radius over 1e8 draws at C2's and C4's radius laws.
* The rejection branch's functions over their argument ranges at C3.

Each test prints the number of differing results and the largest difference
in units in the last place.
"""
import numpy as np
import pytest

from film_grain_amd import capi

pytestmark = pytest.mark.gpu

CHUNK = 1 << 26


def _ulps(a, b):
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    return np.abs(ia - ib)


def _run(ctx, oracle, fn, x):
    import torch
    xd = torch.from_numpy(x).to("cuda:0")
    out = torch.empty_like(xd)
    ctx.math_device(fn, xd, out)
    got = out.cpu().numpy()
    want = oracle.glibc_math(fn, x)
    d = _ulps(got, want)
    return int(np.count_nonzero(d)), int(d.max(initial=0)), x[d != 0][:5]


def test_knuth_threshold_exp_every_f32_below_12(hip_ctx, oracle):
    hi = int(np.array([12.0], np.float32).view(np.uint32)[0])
    bad, worst, ex = 0, 0, []
    for a in range(1, hi, CHUNK):
        bits = np.arange(a, min(a + CHUNK, hi), dtype=np.uint32)
        x = bits.view(np.float32).astype(np.float64)
        n, w, e = _run(hip_ctx, oracle, capi.MATH_EXP_NEG, x)
        bad += n
        worst = max(worst, w)
        ex += list(e)
    print(f"[math] exp(-x), {hi - 1} f32 inputs in (0, 12): {bad} differ, max {worst} ulp, e.g. {ex[:5]}",
          flush=True)
    assert bad == 0


@pytest.mark.parametrize("law", ["c2", "c4_wide", "c4_narrow"])
def test_lognormal_radius_f32_over_1e8_draws(hip_ctx, oracle, law):
    mean, sd = {"c2": (0.1, 0.05), "c4_wide": (0.3, 0.15), "c4_narrow": (0.05, 0.01)}[law]
    d = oracle.derive(mean, sd, "lognorm", 1.0, 0.8, 16)
    mu, sig = float(np.float32(d.log_mu)), float(np.float32(d.log_sigma))
    rng = np.random.default_rng(7)
    bad, raw_bad, worst = 0, 0, 0
    total = 0
    while total < 100_000_000:
        z = rng.standard_normal(CHUNK // 2)
        t = mu + sig * z
        n, w, _ = _run(hip_ctx, oracle, capi.MATH_EXP_F32, t)
        m, w2, _ = _run(hip_ctx, oracle, capi.MATH_EXP, t)
        bad += n
        raw_bad += m
        worst = max(worst, w2)
        total += t.size
    print(f"[math] radius {law}: {total} draws, f32 radius differs {bad}, f64 exp differs {raw_bad} "
          f"(max {worst} ulp)", flush=True)
    assert bad == 0
    assert worst <= 1 and raw_bad <= 1e-3 * total


def test_rejection_branch_math(hip_ctx, oracle):
    rng = np.random.default_rng(11)
    lo = int(np.array([12.0], np.float32).view(np.uint32)[0])
    hi = int(np.array([1000.0], np.float32).view(np.uint32)[0])
    lam = np.arange(lo, hi, dtype=np.uint32).view(np.float32).astype(np.float64)
    report = {}
    report["log(lambda)"] = _run(hip_ctx, oracle, capi.MATH_LOG, lam)[:2]
    report["sqrt(2 lambda)"] = _run(hip_ctx, oracle, capi.MATH_SQRT, 2.0 * lam)[:2]
    report["log_gamma(1 + lambda)"] = _run(hip_ctx, oracle, capi.MATH_LOG_GAMMA, 1.0 + lam)[:2]
    u = (rng.integers(0, 1 << 53, CHUNK, dtype=np.int64).astype(np.float64)) * 2.0 ** -53
    report["tan(pi u)"] = _run(hip_ctx, oracle, capi.MATH_TAN_PI, u)[:2]
    res = np.floor(rng.uniform(0.0, 2000.0, CHUNK))
    report["log_gamma(1 + k)"] = _run(hip_ctx, oracle, capi.MATH_LOG_GAMMA, 1.0 + res)[:2]
    e = -rng.exponential(3.0, CHUNK)
    report["exp(check)"] = _run(hip_ctx, oracle, capi.MATH_EXP, e)[:2]
    for k, (n, w) in report.items():
        print(f"[math] {k}: {n} differ, max {w} ulp", flush=True)
    # exp is fg_exp_dd: glibc's but for its rare misroundings
    assert report["exp(check)"][1] <= 1 and report["exp(check)"][0] <= 1e-3 * e.size
    assert report["sqrt(2 lambda)"][0] == 0
    # log and tan are the device library's: within one unit of glibc; the
    # counts they feed are compared in full (tests/test_gpu_fullsize.py: every
    # pixel of the 4096^2 C3 plane)
    assert report["log(lambda)"][1] <= 1 and report["tan(pi u)"][1] <= 1
    assert report["log_gamma(1 + lambda)"][1] <= 8 and report["log_gamma(1 + k)"][1] <= 8
