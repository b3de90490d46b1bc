"""Seeded random planes across the parameter space (zoom 0.3-4, radius
0.02-0.5, constant or lognormal radii, N from 1 to 2048, cell sizes, filter
sigma, seeds, odd sizes): every formulation of both integrators against the
CPU oracle, bit for bit, and the adaptive renderer at a uniform budget
against the fixed render.  The exactness arguments (the widened sample runs,
the window-test skip, the candidate boxes) are exercised away from the
benchmark's parameters."""
import numpy as np
import pytest

from film_grain_amd import capi, synth
from tests import fgcases

pytestmark = pytest.mark.gpu


def _cases(O, n_cases, seed):
    rng = np.random.default_rng(seed)
    for i in range(n_cases):
        w, h = int(rng.integers(5, 41)), int(rng.integers(5, 41))
        zoom = float(rng.choice([0.3, 0.5, 1.0, 1.3, 2.0, 2.5, 4.0]))
        r = float(np.exp(rng.uniform(np.log(0.02), np.log(0.5))))
        lognorm = bool(rng.integers(0, 2))
        sd = float(rng.uniform(0.1, 0.8)) * r if lognorm else 0.0
        n = int(rng.choice([1, 3, 16, 31, 64, 100, 256, 513, 1024, 2048]))
        # the CPU oracle's work: at most 2e6 (output pixel, sample) points
        while n > 1 and max(1, round(w * zoom)) * max(1, round(h * zoom)) * n > 2_000_000:
            n //= 2
        sigma = float(rng.choice([0.4, 0.8, 1.5]))
        cell = None if rng.random() < 0.6 else float(r * rng.uniform(0.5, 2.0))
        u = synth.sinusoid_field(w, h, int(rng.integers(0, 1 << 30)), n_waves=3)
        if rng.random() < 0.3:
            u = np.clip(u * 1.2 - 0.1, 0.0, 0.999).astype(np.float32)
        name = f"r{i}_{w}x{h}_z{zoom}_r{r:.3f}{'ln' if lognorm else ''}_N{n}_s{sigma}"
        yield fgcases.build_case(O, name, u, r, sd, "lognorm" if lognorm else "const", zoom=zoom,
                                 n_samples=n, sigma=sigma, seed=int(rng.integers(0, 1 << 40)), cell_delta=cell)


def _render(ctx, case, algo, cfg):
    ctx.set_launch_config(**cfg)
    try:
        offs = case.offsets_input if algo == "pixel" else case.offsets
        return ctx.render_host(algo, case.desc, case.lam, offs)
    finally:
        ctx.reset_launch_config()


@pytest.mark.parametrize("seed", [101, 202])
def test_random_planes_pixelwise(hip_ctx, oracle, seed):
    for case in _cases(oracle, 12, seed):
        want = oracle.render_pixelwise(case.plane, case.lam, case.offsets_input)
        variants = [dict(pix_variant=3)]
        if case.n_samples <= 256:
            variants += [dict(pix_variant=1), dict(pix_variant=2)]
        for cfg in variants:
            try:
                got = _render(hip_ctx, case, "pixel", cfg)
            except capi.FilmGrainError as e:
                # the gather formulation refuses a cell window larger than its LDS
                assert e.code == capi.FG_ENEIGHBOURHOOD and cfg["pix_variant"] == 1, (case.name, cfg, e)
                print(f"[random] {case.name} pixel {cfg}: refused ({e})", flush=True)
                continue
            bad = int(np.sum(got != want))
            print(f"[random] {case.name} pixel {cfg}: {bad} differ", flush=True)
            assert bad == 0, (case.name, cfg)


@pytest.mark.parametrize("seed", [303, 404])
def test_random_planes_grainwise(hip_ctx, oracle, seed):
    for case in _cases(oracle, 12, seed):
        want = oracle.render_grainwise(case.plane, case.lam, case.offsets)
        variants = [dict(grain_variant=2)]
        if case.n_samples <= 256:
            variants.append(dict(grain_variant=1))
        for cfg in variants:
            got = _render(hip_ctx, case, "grain", cfg)
            bad = int(np.sum(got != want))
            print(f"[random] {case.name} grain {cfg}: {bad} differ", flush=True)
            assert bad == 0, (case.name, cfg)


def test_random_planes_adaptive_uniform_budget(hip_ctx, oracle):
    import torch
    for case in _cases(oracle, 10, 505):
        n = case.n_samples
        b = max(1, n // 3)
        lam = torch.from_numpy(case.lam).to("cuda:0")
        offs = torch.from_numpy(case.offsets_input).to("cuda:0")
        budget = torch.full((case.desc.out_h, case.desc.out_w), b, dtype=torch.int16, device="cuda:0")
        out = torch.empty((case.desc.out_h, case.desc.out_w), dtype=torch.float32, device="cuda:0")
        hip_ctx.render_adaptive_device(case.desc, lam, offs, budget, out)
        hip_ctx.stream_status()
        p = oracle.plane(case.derived, b, case.seed, case.zoom, case.dist, case.radius)
        want = oracle.render_pixelwise(p, case.lam, np.ascontiguousarray(case.offsets_input[:b]))
        bad = int(np.sum(out.cpu().numpy() != want))
        print(f"[random] {case.name} adaptive budget {b}: {bad} differ", flush=True)
        assert bad == 0, case.name


@pytest.mark.parametrize("r_out", [0.01, 0.05, 0.12, 0.167, 0.1675, 0.18])
def test_grainwise_candidate_lists_near_the_switch(hip_ctx, oracle, r_out):
    """Constant output radii on both sides of the switch to per-cell
    candidate lists (2 R + 1/64 <= 0.35 output pixels, so R <= 0.1671875):
    the sorted kernel gives the oracle's planes bit for bit, for sample
    counts from 1 to 2048 and several zooms."""
    rng = np.random.default_rng(int(r_out * 1e4))
    for zoom, n in [(0.5, 2048), (1.0, 1), (1.7, 100), (3.0, 513)]:
        w, h = 24, 19
        # the oracle walks every (grain, sample) pair: about 1.4 / (pi r^2)
        # grains per input pixel at the field's largest grey level
        grains = w * h * 1.4 / (np.pi * (r_out / zoom) ** 2)
        while n > 1 and (round(w * zoom) * round(h * zoom) * n > 2_000_000 or grains * n > 1e9):
            n //= 2
        u = synth.sinusoid_field(w, h, int(rng.integers(0, 1 << 30)), n_waves=3)
        case = fgcases.build_case(oracle, f"cl_R{r_out}_z{zoom}_N{n}", u, r_out / zoom, 0.0, "const", zoom=zoom,
                                  n_samples=n, sigma=0.8, seed=int(rng.integers(0, 1 << 40)))
        want = oracle.render_grainwise(case.plane, case.lam, case.offsets)
        got = _render(hip_ctx, case, "grain", dict(grain_variant=2))
        bad = int(np.sum(got != want))
        print(f"[random] {case.name} grain sorted: {bad} differ", flush=True)
        assert bad == 0, case.name
