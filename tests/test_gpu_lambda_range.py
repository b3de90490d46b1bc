"""Lambda values outside the accepted range, on MI355X, through the C ABI.

The boundary accepts what lambda_plane produces: values up to MAX_LAMBDA =
1e6 (src/model.rs:11, 262).  Values <= 0 are skipped as the reference skips
them (src/pixelwise.rs:73, src/grainwise.rs:41).  NaN, +inf and values above
1e6 are skipped as well (no Poisson draw runs on them, so no kernel can spin
on a non-finite rejection loop), and the render reports FG_EINVAL.  The rest
of the plane must equal the oracle's render of the same plane with those
values set to 0, for every device formulation of both integrators.
"""
import numpy as np
import pytest

from film_grain_amd import capi, synth
from tests import fgcases

pytestmark = pytest.mark.gpu

BAD = [np.nan, np.inf, 2.0e6, 1.0e30, -np.inf]


def _planted(case):
    lam = case.lam.copy()
    clean = case.lam.copy()
    h, w = lam.shape
    rng = np.random.default_rng(17)
    for v in BAD:
        y, x = int(rng.integers(0, h)), int(rng.integers(0, w))
        lam[y, x] = v
        clean[y, x] = 0.0
    return lam, clean


@pytest.mark.parametrize("algo,variant", [("pixel", 3), ("pixel", 2), ("pixel", 1), ("grain", 2), ("grain", 1)])
def test_out_of_range_lambda_is_skipped_and_reported(hip_ctx, oracle, algo, variant):
    import torch
    u = synth.sinusoid_field(40, 28, 21, n_waves=3)
    case = fgcases.build_case(oracle, f"lam_range_{algo}{variant}", u, 0.1, 0.03, "lognorm", zoom=1.5,
                              n_samples=48, seed=23)
    lam, clean = _planted(case)
    offs = case.offsets_input if algo == "pixel" else case.offsets
    if algo == "pixel":
        want = oracle.render_pixelwise(case.plane, clean, offs)
    else:
        want = oracle.render_grainwise(case.plane, clean, offs)
    dev = torch.device("cuda:0")
    lam_d = torch.from_numpy(lam).to(dev)
    offs_d = torch.from_numpy(np.ascontiguousarray(offs)).to(dev)
    out = torch.full((case.desc.out_h, case.desc.out_w), -1.0, dtype=torch.float32, device=dev)
    key = "pix_variant" if algo == "pixel" else "grain_variant"
    hip_ctx.set_launch_config(**{key: variant})
    try:
        print(f"[lambda range] {algo} variant {variant}", flush=True)
        hip_ctx.render_device(algo, case.desc, lam_d, offs_d, out)
        with pytest.raises(capi.FilmGrainError) as err:
            hip_ctx.stream_status()
        assert err.value.code == capi.FG_EINVAL, str(err.value)
        assert "1e6" in str(err.value)
    finally:
        hip_ctx.reset_launch_config()
    got = out.cpu().numpy()
    assert np.array_equal(got, want), f"max |d| = {np.abs(got - want).max()}"
    # the status word was cleared: a clean render reports nothing
    hip_ctx.render_device(algo, case.desc, torch.from_numpy(clean).to(dev), offs_d, out)
    hip_ctx.stream_status()


def test_host_entry_reports_out_of_range_lambda(hip_ctx, oracle):
    u = synth.sinusoid_field(24, 20, 5, n_waves=2)
    case = fgcases.build_case(oracle, "lam_range_host", u, 0.12, n_samples=16, seed=2)
    lam, _ = _planted(case)
    with pytest.raises(capi.FilmGrainError) as err:
        hip_ctx.render_host("pixel", case.desc, lam, case.offsets_input)
    assert err.value.code == capi.FG_EINVAL
