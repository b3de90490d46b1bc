"""Cell-level parity of the device generators with the oracle, on MI355X.

The north star's bar for the generators: RNG states and Poisson counts per
cell bit-exact, grain geometry within 1 ulp.  Pixel-wise
(src/pixelwise.rs:66-100): the HBM path's grain buffer (fg_pixelwise_grains)
holds, per cell, the Poisson count and every grain's centre and clamped
radius; every cell of the plane's extended grid is compared with the
oracle's cell_rng stream (fgo_cell_count, fgo_cell_dump).  Grain-wise
(src/grainwise.rs:37-58): per input pixel the count and the Xoshiro state
after the count draws (fg_grainwise_pixel_state); the grains the oracle draws
from that state must be the reference's grains (fgo_pixel_dump), which pins
the state itself.
"""
import numpy as np
import pytest

from film_grain_amd import synth
from tests import cellcheck, fgcases

pytestmark = pytest.mark.gpu


def _pixel_cases(O):
    yield fgcases.c2_case(O, w=48, h=32, n_samples=256, channel=0)
    yield fgcases.c1_case(O, size=40, n_samples=64)
    yield fgcases.build_case(O, "rejection_branch", np.full((6, 7), 0.99, np.float32), 0.05, n_samples=16,
                             seed=12, cell_delta=1.0)
    u = synth.sinusoid_field(20, 16, 6, n_waves=3)
    yield fgcases.build_case(O, "lognorm_wide_zoom1.5", u, 0.15, 0.15, "lognorm", zoom=1.5, n_samples=32,
                             seed=99)


def test_pixelwise_cells_equal_oracle(hip_ctx, oracle):
    import torch
    dev = torch.device("cuda:0")
    for case in _pixel_cases(oracle):
        print(f"[cells] {case.name}", flush=True)
        lam = torch.from_numpy(case.lam).to(dev)
        offs = torch.from_numpy(np.ascontiguousarray(case.offsets_input)).to(dev)
        rec, origin, size = hip_ctx.pixelwise_grains(case.desc, lam, offs)
        cells = cellcheck.group_cells(rec, origin, size)
        n_cells, n_grains, inexact, worst = cellcheck.compare_cells(oracle, case.plane, case.lam, cells, origin, size)
        print(f"[cells] {case.name}: grid {size}, {n_cells} nonempty cells, {n_grains} grains, "
              f"{inexact} not bit-equal, max {worst} ulp", flush=True)
        assert n_grains == len(rec) and n_cells > 0
        assert worst <= 1
        assert inexact == 0, "grain geometry bit-exact on these cases"


@pytest.mark.parametrize("which", ["c3_40", "lognorm_zoom2"])
def test_grainwise_pixel_streams_equal_oracle(hip_ctx, oracle, which):
    import torch
    if which == "c3_40":
        case = fgcases.c3_case(oracle, size=40, n_samples=128)
    else:
        u = synth.sinusoid_field(20, 14, 3, n_waves=4)
        case = fgcases.build_case(oracle, which, u, 0.2, 0.1, "lognorm", zoom=2.0, n_samples=32, seed=3)
    lam = torch.from_numpy(case.lam).to(torch.device("cuda:0"))
    q, st = hip_ctx.grainwise_pixel_state(case.desc, lam)
    n_grains = 0
    for y in range(case.desc.in_h):
        for x in range(case.desc.in_w):
            h, s0, qo, want = oracle.pixel_dump(case.plane, case.lam, x, y, max_grains=4096)
            assert int(q[y, x]) == qo, f"pixel ({x}, {y}): count {q[y, x]} vs {qo}"
            if qo == 0:
                continue
            got = cellcheck.replay_pixel_grains(oracle, case.plane, x, y, st[y, x], qo)
            assert np.array_equal(got, want), f"pixel ({x}, {y}): grains differ"
            n_grains += qo
    print(f"[pixel streams] {which}: {n_grains} grains replayed", flush=True)
    assert n_grains > 0
