"""The work counters behind the roofline figures, against the oracle's.

The pixel-wise FP32 FLOP count is 16 + 6 T per (pixel, sample), with T the
disk tests the reference performs in its own loop order with its early exit
(src/pixelwise.rs:47-106).  The grain-wise count is 8 per (grain, sample) pair
plus 5 per disk test (src/grainwise.rs:46-104).  The device counts them with
its instrumented twins (fg_pixelwise_stats_device, fg_grainwise_stats_device);
the oracle counts them in its restatement of the reference loops.  Both must
agree exactly, and the instrumented renders must equal the oracle's planes.
"""
import numpy as np
import pytest

from film_grain_amd import synth
from tests import fgcases

pytestmark = pytest.mark.gpu


def _device_stats(ctx, algo, case):
    import torch
    dev = torch.device("cuda:0")
    lam = torch.from_numpy(case.lam).to(dev)
    offs = case.offsets_input if algo == "pixel" else case.offsets
    offs_d = torch.from_numpy(np.ascontiguousarray(offs)).to(dev)
    out = torch.empty((case.desc.out_h, case.desc.out_w), dtype=torch.float32, device=dev)
    cnt = ctx.stats_device(algo, case.desc, lam, offs_d, out)
    return cnt, out.cpu().numpy()


@pytest.mark.parametrize("which", ["c1_256", "c2_96x64", "lognorm_zoom1.5"])
def test_pixelwise_T_equals_oracle(hip_ctx, oracle, which):
    if which == "c1_256":
        case = fgcases.c1_case(oracle, size=256, n_samples=64)
    elif which == "c2_96x64":
        case = fgcases.c2_case(oracle, w=96, h=64, n_samples=256, channel=2)
    else:
        u = synth.sinusoid_field(40, 30, 8, n_waves=3)
        case = fgcases.build_case(oracle, which, u, 0.15, 0.1, "lognorm", zoom=1.5, n_samples=40, seed=8)
    want, c = oracle.render_pixelwise(case.plane, case.lam, case.offsets_input, counters=True)
    (points, tests, hits), got = _device_stats(hip_ctx, "pixel", case)
    print(f"[{which}] points {points} T {tests} hits {hits}; oracle {c.points} {c.tests} {c.hits} "
          f"T/point {tests / max(points, 1):.4f}", flush=True)
    assert np.array_equal(got, want)
    assert (points, tests, hits) == (c.points, c.tests, c.hits)


@pytest.mark.parametrize("which", ["c3_96", "lognorm_zoom2"])
def test_grainwise_counters_equal_oracle(hip_ctx, oracle, which):
    if which == "c3_96":
        case = fgcases.c3_case(oracle, size=96, n_samples=128)
    else:
        u = synth.sinusoid_field(36, 28, 3, n_waves=4)
        case = fgcases.build_case(oracle, which, u, 0.2, 0.1, "lognorm", zoom=2.0, n_samples=96, seed=3)
    want, c = oracle.render_grainwise_counted(case.plane, case.lam, case.offsets)
    (grains, pairs, tests, hits), got = _device_stats(hip_ctx, "grain", case)
    print(f"[{which}] grains {grains} pairs {pairs} tests {tests} hits {hits}; oracle {c}", flush=True)
    assert np.array_equal(got, want)
    assert (grains, pairs, tests, hits) == (c["grains"], c["pairs"], c["tests"], c["hits"])
