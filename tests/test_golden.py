"""The committed golden fixtures (tests/golden/fixtures.npz) on the CPU.

The reference has no tests or fixtures; these vectors were written by the
oracle (tests/golden/make_golden.py).  Here the oracle must reproduce every
one of them exactly, and the per-cell hashes and seeded states must equal
their closed forms (src/rng.rs:26-52: mix3 of the cell stream, then
seed_from_u64's four splitmix64 outputs) computed with Python integers,
independently of the oracle.
"""
import numpy as np
import pytest

from tests.golden import fixture

MASK = (1 << 64) - 1
PHI = 0x9E3779B97F4A7C15
CELL_STREAM = 0xA24B1C30BEBCCF59


def _sm(x):
    x = (x + PHI) & MASK
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & MASK


def _mix3(seed, stream, a, b):
    s = seed ^ stream
    s = _sm(_rotl((s + (a & MASK)) & MASK, 17))
    return _sm(_rotl((s + (b & MASK)) & MASK, 41))


@pytest.fixture(scope="module")
def fx():
    return fixture.load()


@pytest.mark.parametrize("prefix", ["c1", "c2"])
def test_cell_fixtures(oracle, fx, prefix):
    p = fixture.plane(oracle, fx, prefix)
    lam = fx[f"{prefix}_lam"]
    i0, j0 = (int(v) for v in fx[f"{prefix}_cells_origin"])
    q = fx[f"{prefix}_cells_q"]
    assert q.sum() > 0
    for dj in range(16):
        for di in range(16):
            i, j = i0 + di, j0 + dj
            h = _mix3(p.seed, CELL_STREAM, i, j)
            assert int(fx[f"{prefix}_cells_hash"][dj, di]) == h
            assert [int(v) for v in fx[f"{prefix}_cells_state"][dj, di]] == [_sm((h + k * PHI) & MASK)
                                                                              for k in range(4)]
            h2, s2, q2, g2 = oracle.cell_dump(p, lam, i, j, max_grains=8)
            assert (h2, q2) == (h, int(q[dj, di]))
            assert np.array_equal(g2, fx[f"{prefix}_cells_grains"][dj, di, :len(g2)])


def test_c1_patch_fixture(oracle, fx):
    p = fixture.plane(oracle, fx, "c1")
    region = tuple(int(v) for v in fx["c1_patch_region"])
    got = oracle.render_pixelwise(p, fx["c1_lam"], fx["c1_offsets_input"], region=region)
    assert np.array_equal(got, fx["c1_patch"])


def test_c3_bits_fixture(oracle, fx):
    p = fixture.plane(oracle, fx, "c3")
    out, bits, pairs = oracle.render_grainwise(p, fx["c3_lam"], fx["c3_offsets"], want_bitsets=True)
    x0, y0 = (int(v) for v in fx["c3_bits_origin"])
    assert np.array_equal(bits[y0:y0 + 16, x0:x0 + 16], fx["c3_bits"])
    assert np.array_equal(out[y0:y0 + 16, x0:x0 + 16], fx["c3_patch"])
    assert pairs == int(fx["c3_pairs"][0])
    # the patch's output is its bitsets' popcount over N
    pop = np.array([[sum(bin(int(w)).count("1") for w in px) for px in row] for row in fx["c3_bits"]])
    assert np.array_equal(fx["c3_patch"], (pop.astype(np.float32) * np.float32(1.0 / 128)).astype(np.float32))
