"""Pins the CPU oracle (oracle/fg_oracle.c) before anything is checked against it.

The reference has no tests, golden vectors or fixtures of its own, and its
third-party random stack (rand 0.8.5, rand_distr 0.4.3, rand_chacha 0.3.1,
statrs 0.16.1) cannot be built here.  Each oracle function is therefore pinned
by a published known-answer vector, a closed form, a statistical invariant or a
brute-force recomputation, chosen so that a plausible slip (wrong constant,
rotation, sign extension, operation order, branch) fails.
"""
import math

import numpy as np
import pytest
from scipy import special, stats

MASK = (1 << 64) - 1
PHI = 0x9E3779B97F4A7C15


# ---------------------------------------------------------------- splitmix64 / mix3
def test_splitmix64_published_sequence(oracle):
    # Vigna's splitmix64.c started from x = 0: the first four outputs.
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    got = [oracle.lib().fgo_splitmix64((k * PHI) & MASK) for k in range(4)]
    assert got == want


def _sm(x):
    x = (x + PHI) & MASK
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & MASK


def test_mix3_closed_form_with_negative_indices(oracle):
    # src/rng.rs:40-44 with i32 -> i64 -> u64 sign extension; the python big-int
    # statement catches a zero-extension or a rotate/splitmix order slip.
    for seed in (0, 7, 5489, 0xFFFFFFFFFFFFFFFF):
        for stream in (0xA24B1C30BEBCCF59, 0x6935FA5C55F65F1B):
            for a, b in ((0, 0), (-1, 3), (38400, -17), (-2147483648, 2147483647)):
                s = seed ^ stream
                s = _sm(_rotl((s + (a & MASK)) & MASK, 17))
                want = _sm(_rotl((s + (b & MASK)) & MASK, 41))
                assert oracle.lib().fgo_mix3(seed, stream, a, b) == want


# ---------------------------------------------------------------- xoshiro256++
def test_xoshiro256pp_reference_outputs(oracle):
    # rand_xoshiro's reference vector for Xoshiro256PlusPlus from state [1, 2, 3, 4].
    want = [41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205,
            9973669472204895162, 14011001112246962877, 12406186145184390807,
            15849039046786891736, 10450023813501588000]
    x = oracle.Xoshiro(state=[1, 2, 3, 4])
    assert [x.next_u64() for _ in range(10)] == want


def test_xoshiro_seed_is_splitmix_sequence(oracle):
    for h in (0, 1, 0xDEADBEEF, MASK):
        x = oracle.Xoshiro(seed=h)
        assert [int(v) for v in x.s] == [_sm((h + k * PHI) & MASK) for k in range(4)]


def test_next_u32_is_high_half(oracle):
    a, b = oracle.Xoshiro(seed=99), oracle.Xoshiro(seed=99)
    for _ in range(16):
        assert a.next_u32() == b.next_u64() >> 32


# ---------------------------------------------------------------- uniform / standard
@pytest.mark.parametrize("delta", [1 / 9, 0.1, 0.05, 0.25, 1.0, 1 / 7, 0.37])
def test_uniform_f32_range_exhaustive(oracle, delta):
    d = np.float32(delta)
    scale = np.float32(oracle.lib().fgo_uniform_f32_scale(0.0, d))
    m = np.arange(1 << 23, dtype=np.uint32)
    v01 = (m | np.uint32(0x3F800000)).view(np.float32) - np.float32(1.0)
    out = v01 * scale + np.float32(0.0)
    assert out.min() == 0.0 and out.max() < d
    # the shrink loop stops at the largest admissible scale
    up = np.nextafter(scale, np.float32(np.inf))
    top = np.float32(v01[-1] * up) + np.float32(0.0)
    assert up >= d or top >= d


def test_standard_and_open01_ranges(oracle):
    x = oracle.Xoshiro(seed=5)
    s = np.array([x.standard_f64() for _ in range(20000)])
    o = np.array([x.open01_f64() for _ in range(20000)])
    assert s.min() >= 0.0 and s.max() < 1.0 and abs(s.mean() - 0.5) < 0.01
    assert o.min() > 0.0 and o.max() < 1.0 and abs(o.mean() - 0.5) < 0.01


# ---------------------------------------------------------------- log_gamma / Poisson
def test_log_gamma_against_libm(oracle):
    for x in [1.0, 1.5, 2.0, 3.7, 10.0, 12.5, 100.0, 880.0, 1e4]:
        got = oracle.lib().fgo_log_gamma(x)
        assert abs(got - math.lgamma(x)) <= 2e-10 * max(1.0, abs(math.lgamma(x)))


def _poisson_draws(oracle, lam, n, seed):
    x = oracle.Xoshiro(seed=seed)
    return np.array([x.poisson(lam) for _ in range(n)])


@pytest.mark.parametrize("lam", [0.01, 0.176, 1.76, 5.0, 11.9, 12.0, 100.0, 880.0])
def test_poisson_moments_and_chisquare(oracle, lam):
    n = 40000
    d = _poisson_draws(oracle, lam, n, seed=int(lam * 1000) + 1)
    assert np.all(d == np.floor(d)) and d.min() >= 0
    se = math.sqrt(lam / n)
    assert abs(d.mean() - lam) < 5 * se
    assert abs(d.var() - lam) < 6 * lam * math.sqrt(2.0 / n) + 1e-3
    lo, hi = stats.poisson.ppf(0.001, lam), stats.poisson.ppf(0.999, lam)
    edges = np.unique(np.linspace(lo, hi + 1, min(30, int(hi - lo) + 2)).astype(int))
    if len(edges) >= 3:
        obs = np.histogram(d, bins=np.append(edges, np.inf))[0][:-1]
        cdf = stats.poisson.cdf(edges - 1, lam)
        exp = np.diff(cdf) * n
        keep = exp > 5
        chi2 = ((obs[:-1][keep[:-1]] - exp[:-1][keep[:-1]]) ** 2 / exp[:-1][keep[:-1]]).sum()
        dof = max(1, keep[:-1].sum() - 1)
        assert stats.chi2.sf(chi2, dof) > 1e-4


def test_poisson_tiny_lambda_is_zero(oracle):
    # exp(-lambda) rounds to 1.0, Knuth's loop never runs and the sample is
    # 0 - 1 = -1.0, which the integrators' `as u32` saturates to 0.
    d = _poisson_draws(oracle, 1e-18, 100, seed=3)
    assert np.all(d == -1.0)


# ---------------------------------------------------------------- ziggurat
def test_ziggurat_tables_recurrence_and_literals(oracle):
    X, F = oracle.zig_tables()
    R, V = 3.6541528853610088, 0.00492867323399
    assert X[1] == pytest.approx(R, rel=1e-17) and X[256] == 0.0 and F[256] == 1.0
    assert np.allclose(F, np.exp(-X * X / 2.0), rtol=1e-14, atol=1e-17)
    for i in range(2, 256):
        assert X[i] == pytest.approx(math.sqrt(-2.0 * math.log(V / X[i - 1] + F[i - 1])), rel=1e-13)
    # decimal literals of rand_distr's ziggurat_tables.rs (first entries)
    assert X[0] == 3.910757959537090045 and X[2] == 3.449278298560964462
    assert X[3] == 3.320244733839166074 and X[4] == 3.224575052047029100
    assert F[0] == 0.000477467764586655 and F[1] == 0.001260285930498598
    assert F[2] == 0.002609072746106363


def test_product_ziggurat_header_equals_oracle_tables(oracle):
    import re
    from pathlib import Path
    txt = (Path(__file__).resolve().parents[1] / "film_grain_amd/csrc/fg_zig_tables.h").read_text()
    def table(name):
        body = txt.split(f"#define {name}")[1].split("\n#")[0]
        return np.array([float(v) for v in re.findall(r"[-0-9.]+", body)])
    X, F = oracle.zig_tables()
    assert np.array_equal(table("FG_ZIG_NORM_X_VALUES"), X)
    assert np.array_equal(table("FG_ZIG_NORM_F_VALUES"), F)


def test_standard_normal_distribution(oracle):
    x = oracle.Xoshiro(seed=11)
    z = np.array([x.standard_normal() for _ in range(100000)])
    assert abs(z.mean()) < 0.015 and abs(z.var() - 1.0) < 0.02
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # the tail branch (layer 0) is exercised and lands beyond R
    tail = np.mean(np.abs(z) > 3.6541528853610088)
    assert abs(tail - 2 * stats.norm.sf(3.6541528853610088)) < 5e-4


def test_lognormal_moments(oracle):
    x = oracle.Xoshiro(seed=12)
    mu, sig = -2.4141569, 0.4723807
    r = np.array([x.lognormal(mu, sig) for _ in range(50000)])
    assert abs(r.mean() - 0.1) < 0.002 and abs(r.std() - 0.05) < 0.002


# ---------------------------------------------------------------- ChaCha / StdRng
def test_chacha20_rfc7539_keystream(oracle):
    # RFC 7539 appendix A.1 test vectors #1 and #2 (zero key and nonce, block
    # counters 0 and 1), read as little-endian 32-bit words.
    b0 = [0xade0b876, 0x903df1a0, 0xe56a5d40, 0x28bd8653, 0xb819d2bd, 0x1aed8da0, 0xccef36a8,
          0xc70d778b, 0x7c5941da, 0x8d485751, 0x3fe02477, 0x374ad8b8, 0xf4b8436a, 0x1ca11815,
          0x69b687c3, 0x8665eeb2]
    b1 = [0xbee7079f, 0x7a385155, 0x7c97ba98, 0x0d082d73, 0xa0290fcb, 0x6965e348, 0x3e53c612,
          0xed7aee32, 0x7621b729, 0x434ee69c, 0xb03371d5, 0xd539d874, 0x281fed31, 0x45fb0a51,
          0x1f0ae1ac, 0x6f4d794b]
    assert list(oracle.chacha_block([0] * 8, 0, 0, 20)) == b0
    assert list(oracle.chacha_block([0] * 8, 1, 0, 20)) == b1


def test_stdrng_words_follow_block_order(oracle):
    # StdRng.next_u64 reads consecutive 32-bit words of 4 consecutive blocks.
    seed = 12345
    state = seed
    key = []
    for _ in range(8):
        state = (state * 6364136223846793005 + 11634580027462260723) & MASK
        xs = (((state >> 18) ^ state) >> 27) & 0xFFFFFFFF
        rot = state >> 59
        key.append(((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF)
    words = np.concatenate([oracle.chacha_block(key, c, 0, 12) for c in range(8)]).astype(np.uint64)
    want = words[0::2] | (words[1::2] << np.uint64(32))
    assert np.array_equal(oracle.stdrng_u64(seed, 64), want)


def test_make_offsets_statistics(oracle):
    o = oracle.make_offsets(7, 4096, 0.8).astype(np.float64)
    assert abs(o.mean()) < 0.05 and abs(o.std() - 0.8) < 0.03
    assert stats.kstest(o[:, 0] / 0.8, "norm").pvalue > 1e-3


# ---------------------------------------------------------------- parameters
def test_inverse_normal_cdf_published_quantiles(oracle):
    L = oracle.lib()
    assert L.fgo_inverse_normal_cdf(0.975) == pytest.approx(1.959963984540054, abs=1e-14)
    assert L.fgo_inverse_normal_cdf(0.999) == pytest.approx(3.090232306167813, abs=1e-14)
    assert L.fgo_inverse_normal_cdf(0.5) == 0.0
    for p in np.linspace(1e-9, 1 - 1e-9, 41):
        assert special.ndtr(L.fgo_inverse_normal_cdf(p)) == pytest.approx(p, rel=1e-12, abs=1e-17)


def test_derived_parameters_of_the_benchmark_configs(oracle):
    d = oracle.derive(0.1, 0.05, "lognorm", 2.0, 0.8, 256, seed=7, in_w=1920, in_h=1080)
    assert np.float32(d.rm) == np.float32(0.385047) or abs(d.rm - 0.385047) < 1e-6
    assert abs(d.log_mu - -2.4141569) < 1e-6 and abs(d.log_sigma - 0.4723807) < 1e-6
    assert d.delta == np.float32(0.1) and (d.out_w, d.out_h) == (3840, 2160)
    assert abs(d.inv_e_pi_r2 - 25.46479) < 1e-4
    d = oracle.derive(0.12, 0.0, "const", 1.0, 0.8, 64, seed=42, in_w=256, in_h=256)
    assert d.delta == np.float32(1.0) / np.float32(9.0) and d.rm == np.float32(0.12)
    assert abs(d.inv_e_pi_r2 - 22.10485) < 1e-4
    d = oracle.derive(0.05, 0.0, "const", 1.0, 0.8, 128, seed=11, in_w=4096, in_h=4096)
    assert abs(d.inv_e_pi_r2 - 127.32395) < 1e-3


def test_choose_algorithm_routes_the_configs(oracle):
    L = oracle.lib()
    assert L.fgo_choose_algorithm(0.12, 0.0, 0.12, 64, 0) == 2   # C1 -> pixel
    assert L.fgo_choose_algorithm(0.1, 0.05, 0.385047, 256, 0) == 1  # C2 -> grain (forced pixel)
    assert L.fgo_choose_algorithm(0.05, 0.0, 0.05, 128, 0) == 1  # C3 -> grain
    assert L.fgo_choose_algorithm(0.3, 0.0, 0.3, 24, 0) == 2     # small N, tight rm -> pixel
    assert L.fgo_choose_algorithm(0.1, 0.0, 1.0, 32, 0) == 2     # first rule wins before rm/mean
    assert L.fgo_choose_algorithm(0.1, 0.0, 1.0, 80, 0) == 1     # rm/mean > 8 -> grain
    assert L.fgo_choose_algorithm(0.4, 0.1, 0.8, 20, 0) == 2     # N <= 24 and rm/mean < 5
    assert L.fgo_choose_algorithm(0.4, 0.1, 0.8, 30, 0) == 1     # otherwise grain


def test_lambda_plane_closed_form(oracle):
    u = np.array([0.0, 0.25, 0.5, 0.999, 1.0, 2.0, -1.0, np.nan], dtype=np.float32)
    n, mx = oracle.normalize_plane(u)
    assert mx == 2.0
    lam = oracle.lambda_plane(n, np.float32(25.46479))
    ref = -np.float32(25.46479) * np.log(np.maximum(np.float32(1) - np.clip(n[:-1], 0, np.float32(1 - 1e-6)),
                                                    np.float32(1e-6)))
    assert np.allclose(lam[:-1], ref, rtol=2e-7)
    assert lam[0] == 0.0 and np.isfinite(lam[-1]) and lam[-1] > 0  # NaN -> maximal density


# ---------------------------------------------------------------- integrators
def _brute_pixelwise(oracle, case, region):
    """Every grain of every cell in a wide neighbourhood, tested against every
    (pixel, sample) point with the reference's f32 window and disk test."""
    p, lam = case.plane, case.lam
    f = np.float32
    rm, delta, iz = f(p.rm), f(p.delta), f(1.0) / f(p.zoom)
    x0, y0, x1, y1 = region
    pts = []
    for y in range(y0, y1):
        for x in range(x0, x1):
            for k in range(p.n_samples):
                pts.append(((f(x) + f(0.5)) * iz - case.offsets_input[k, 0],
                            (f(y) + f(0.5)) * iz - case.offsets_input[k, 1]))
    pts = np.array(pts, dtype=np.float32)
    ilo = int(np.floor((pts[:, 0].min() - 2 * rm) / delta)) - 2
    ihi = int(np.floor((pts[:, 0].max() + 2 * rm) / delta)) + 2
    jlo = int(np.floor((pts[:, 1].min() - 2 * rm) / delta)) - 2
    jhi = int(np.floor((pts[:, 1].max() + 2 * rm) / delta)) + 2
    hit = np.zeros(len(pts), dtype=bool)
    i0 = np.floor((pts[:, 0] - rm) / delta).astype(np.int64)
    i1 = np.floor((pts[:, 0] + rm) / delta).astype(np.int64)
    j0 = np.floor((pts[:, 1] - rm) / delta).astype(np.int64)
    j1 = np.floor((pts[:, 1] + rm) / delta).astype(np.int64)
    for i in range(ilo, ihi + 1):
        for j in range(jlo, jhi + 1):
            _, _, q, g = oracle.cell_dump(p, lam, i, j, max_grains=4096)
            if q == 0:
                continue
            inwin = (i0 <= i) & (i <= i1) & (j0 <= j) & (j <= j1)
            for cx, cy, r in g:
                if r <= 0:
                    continue
                dx = pts[:, 0] - f(cx)
                dy = pts[:, 1] - f(cy)
                hit |= inwin & (dx * dx + dy * dy <= f(r) * f(r))
    cnt = hit.reshape(y1 - y0, x1 - x0, p.n_samples).sum(axis=2).astype(np.float32)
    return cnt * (f(1.0) / f(max(p.n_samples, 1)))


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_pixelwise_oracle_equals_brute_force(oracle, which):
    from tests import fgcases
    case = fgcases.c1_case(oracle, size=16, n_samples=8) if which == "c1" else \
        fgcases.c2_case(oracle, w=12, h=8, n_samples=8)
    region = (3, 2, 9, 7)
    want = _brute_pixelwise(oracle, case, region)
    got = oracle.render_pixelwise(case.plane, case.lam, case.offsets_input, region=region)
    assert np.array_equal(got, want)


def test_grainwise_oracle_equals_brute_force(oracle):
    from tests import fgcases
    f = np.float32
    u = fgcases.synth.sinusoid_field(9, 7, 4, n_waves=2)
    case = fgcases.build_case(oracle, "g", u, 0.3, 0.1, "lognorm", zoom=1.5, n_samples=12, seed=4)
    p = case.plane
    bits = np.zeros((p.out_h, p.out_w, p.n_samples), dtype=bool)
    z = f(p.zoom)
    for y in range(p.in_h):
        for x in range(p.in_w):
            _, _, q, g = oracle.pixel_dump(p, case.lam, x, y, max_grains=8192)
            for cx, cy, r in g:
                if r <= 0:
                    continue
                R = f(r) * z
                for k in range(p.n_samples):
                    tx = f(cx) * z + case.offsets[k, 0]
                    ty = f(cy) * z + case.offsets[k, 1]
                    for oy in range(p.out_h):
                        for ox in range(p.out_w):
                            dx = (f(ox) + f(0.5)) - tx
                            dy = (f(oy) + f(0.5)) - ty
                            if dx * dx + dy * dy <= R * R:
                                bits[oy, ox, k] = True
    want = bits.sum(axis=2).astype(np.float32) * (f(1) / f(p.n_samples))
    got = oracle.render_grainwise(p, case.lam, case.offsets)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("u0", [0.2, 0.6])
def test_pixelwise_coverage_invariant(oracle, u0):
    # Boolean model: a flat field u has expected coverage u when no grain is clipped.
    from tests import fgcases
    case = fgcases.build_case(oracle, "flat", np.full((48, 48), u0, np.float32), 0.12, n_samples=32, seed=21)
    out = oracle.render_pixelwise(case.plane, case.lam, case.offsets_input)
    assert abs(float(out[8:-8, 8:-8].mean()) - u0) < 0.01


def test_grainwise_and_pixelwise_agree_in_mean(oracle):
    from tests import fgcases
    case = fgcases.build_case(oracle, "flat", np.full((40, 40), 0.45, np.float32), 0.1, n_samples=32, seed=22)
    a = oracle.render_pixelwise(case.plane, case.lam, case.offsets_input)[6:-6, 6:-6].mean()
    b = oracle.render_grainwise(case.plane, case.lam, case.offsets)[6:-6, 6:-6].mean()
    assert abs(a - 0.45) < 0.015 and abs(b - 0.45) < 0.015


def test_zero_intensity_renders_zero(oracle):
    from tests import fgcases
    case = fgcases.build_case(oracle, "zero", np.zeros((10, 10), np.float32), 0.1, 0.05, "lognorm",
                              zoom=2.0, n_samples=16, seed=1)
    assert oracle.render_pixelwise(case.plane, case.lam, case.offsets_input).max() == 0.0
    assert oracle.render_grainwise(case.plane, case.lam, case.offsets).max() == 0.0
