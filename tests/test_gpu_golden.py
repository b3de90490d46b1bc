"""The committed golden fixtures (tests/golden/fixtures.npz) on MI355X.

The device must reproduce the oracle-written vectors through the C ABI: the
per-cell Poisson counts and grains of the 16 x 16-cell patches at C1 and C2
parameters (from the HBM generation's grain buffer), the C1 32 x 32 output
patch, and the C3 grain-wise patch (its values are the fixture bitsets'
popcounts over N).
"""
import numpy as np
import pytest

from tests import cellcheck
from tests.golden import fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fx():
    return fixture.load()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda:0"))


@pytest.mark.parametrize("prefix", ["c1", "c2"])
def test_device_cells_match_fixture(hip_ctx, fx, prefix):
    desc = fixture.desc(fx, prefix)
    rec, origin, size = hip_ctx.pixelwise_grains(desc, _dev(fx[f"{prefix}_lam"]), _dev(fx[f"{prefix}_offsets_input"]))
    cells = cellcheck.group_cells(rec, origin, size)
    i0, j0 = (int(v) for v in fx[f"{prefix}_cells_origin"])
    q = fx[f"{prefix}_cells_q"]
    for dj in range(16):
        for di in range(16):
            got = cells.get((i0 + di, j0 + dj))
            n = 0 if got is None else len(got)
            assert n == int(q[dj, di]), f"cell ({i0 + di}, {j0 + dj})"
            if n:
                assert np.array_equal(got[:8], fx[f"{prefix}_cells_grains"][dj, di, :min(n, 8)])


def test_device_c1_patch_matches_fixture(hip_ctx, fx):
    desc = fixture.desc(fx, "c1")
    out = hip_ctx.render_host("pixel", desc, fx["c1_lam"], fx["c1_offsets_input"])
    x0, y0, x1, y1 = (int(v) for v in fx["c1_patch_region"])
    assert np.array_equal(out[y0:y1, x0:x1], fx["c1_patch"])


def test_device_c3_patch_matches_fixture(hip_ctx, fx):
    desc = fixture.desc(fx, "c3")
    out = hip_ctx.render_host("grain", desc, fx["c3_lam"], fx["c3_offsets"])
    x0, y0 = (int(v) for v in fx["c3_bits_origin"])
    assert np.array_equal(out[y0:y0 + 16, x0:x0 + 16], fx["c3_patch"])
