"""The splat's candidate lists at their boundaries.  Lists are kept per
radius class (rz < (k + 1)/64 output px), only for narrow grains (rz < 0.5)
and only for tables of at most 2048 samples; everything else walks the
x-run.  Constant radii placed just either side of class edges and of the
narrow limit must give the oracle's planes bit for bit (N = 4096: runs
only), and adaptive renders must give each pixel the fixed render at its
budget."""
import numpy as np
import pytest

from film_grain_amd import synth
from tests import fgcases
from tests.test_gpu_adaptive import _check_against_fixed

pytestmark = pytest.mark.gpu

SLACK = lambda out_side: 4e-3 + 2e-6 * out_side  # noqa: E731  (the splat kernel's widening of rz)


# 1025 and above: past the largest N whose tables hold (word, bit) pairs;
# 4096: past the largest N with lists
@pytest.mark.parametrize("zoom,n", [(2.0, 256), (2.7, 1024), (1.0, 100), (1.5, 1025), (1.0, 2048), (1.0, 4096),
                                    (1.0, 16)])
def test_class_edges_and_narrow_limit(hip_ctx, oracle, zoom, n):
    rng = np.random.default_rng(int(zoom * 100) + n)
    w, h = 22, 17
    out_side = max(round(w * zoom), round(h * zoom))
    # rz just below / above class edges k/64, and around the narrow limit 0.5
    targets = [10 / 64 - 1e-5, 10 / 64 + 1e-5, 21 / 64 - 1e-5, 31 / 64 + 2e-5, 0.5 - 1e-5, 0.5 + 1e-3, 0.61]
    for rz in targets:
        r = (rz - SLACK(out_side)) / zoom
        u = synth.sinusoid_field(w, h, int(rng.integers(0, 1 << 30)), n_waves=3)
        case = fgcases.build_case(oracle, f"edge_rz{rz:.6f}_z{zoom}_N{n}", u, r, 0.0, "const", zoom=zoom,
                                  n_samples=n, sigma=0.8, seed=int(rng.integers(0, 1 << 40)))
        want = oracle.render_pixelwise(case.plane, case.lam, case.offsets_input)
        hip_ctx.set_launch_config(pix_variant=3)
        try:
            got = hip_ctx.render_host("pixel", case.desc, case.lam, case.offsets_input)
        finally:
            hip_ctx.reset_launch_config()
        bad = int(np.sum(got != want))
        print(f"[lists] {case.name}: {bad} differ", flush=True)
        assert bad == 0, case.name


def test_adaptive_levels_with_lists(hip_ctx, oracle):
    """zoom 1, r = 0.1, levels 128 to 1024, each with its own lists."""
    import torch
    w, h = 40, 28
    u = synth.sinusoid_field(w, h, 77, n_waves=3)
    case = fgcases.build_case(oracle, "levels", u, 0.1, 0.05, "lognorm", zoom=1.0, n_samples=1024, sigma=0.8,
                              seed=9)
    rng = np.random.default_rng(5)
    # 8x8-pixel blocks of one budget each, so that tiles take different levels
    blocks = rng.choice([100, 200, 256, 300, 512, 700, 1024], size=(h // 8 + 1, w // 8 + 1))
    budget = np.kron(blocks, np.ones((8, 8), dtype=np.int64))[:case.desc.out_h, :case.desc.out_w]
    _check_against_fixed(hip_ctx, torch, case, budget, "levels")
