"""Pins the oracle's bit-to-float conversions and its PCG32 step.

Each conversion is driven with chosen next_u64 values: a Xoshiro256++ state
(0, s1, s2, rotr(v, 23)) returns v as its next output (result = rotl(s0 + s3,
23) + s0), which the first assertion checks.  The properties asserted follow
from the published algorithms (rand 0.8.5) and fail on plausible slips:

* Standard<f64> (src/rng.rs via rand_distr's Poisson uniforms): the 53 high
  bits times 2^-53.  Monotone in v (a low-bit extraction is not), on the
  2^-53 lattice with odd multiples present (a shift by 12 gives only even
  ones), blind to the 11 low bits and sensitive to bit 11.
* Open01<f64> (the ziggurat tail): a [1, 2) mantissa from the 52 high bits
  minus (1 - 2^-53), so always an odd multiple of 2^-53 in (0, 1), monotone,
  blind to the 12 low bits.
* UniformFloat<f32>::sample (grain centres, src/pixelwise.rs:87-88): the 23
  high bits of next_u32 (the high half of next_u64) as a [1, 2) mantissa,
  minus 1, times scale, plus low.  Blind to the 41 low bits of next_u64 (a
  `next_u64 & 0x7FFFFF` or low-half next_u32 slip is not), on the 2^-23
  lattice at scale 1, monotone.
* PCG32 (rand_core 0.6.4 seed_from_u64 for StdRng, src/rng.rs:15): the
  published pcg32 demo output (seed 42, stream 54).
"""
import ctypes as C

import numpy as np
import pytest

MASK = (1 << 64) - 1


def _rotr(x, k):
    return ((x >> k) | (x << (64 - k))) & MASK


def _state(v):
    return np.array([0, 0x0123456789ABCDEF, 0xFEDCBA9876543210, _rotr(v, 23)], dtype=np.uint64)


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _draw(oracle, fn, v, *args):
    s = _state(v)
    return getattr(oracle.lib(), fn)(_u64p(s), *args)


def _values():
    rng = np.random.default_rng(31)
    vs = {0, 1, 0x7FF, 0x800, 0xFFF, 0x1000, (1 << 41) - 1, 1 << 41, 1 << 63, MASK, MASK - 0x7FF, MASK - 0x800}
    vs |= {int(x) for x in rng.integers(0, 1 << 63, 400, dtype=np.int64)}
    vs |= {int(x) | (1 << 63) for x in rng.integers(0, 1 << 63, 400, dtype=np.int64)}
    return sorted(vs)


def test_chosen_state_returns_chosen_output(oracle):
    for v in _values()[:50]:
        assert _draw(oracle, "fgo_xoshiro_next_u64", v) == v


def test_standard_f64_bits(oracle):
    vs = _values()
    f = np.array([_draw(oracle, "fgo_standard_f64", v) for v in vs])
    assert np.all(np.diff(f) >= 0), "not monotone in the source u64"
    k = f * 2.0 ** 53
    assert np.all(k == np.floor(k)), "off the 2^-53 lattice"
    assert np.any(k % 2 == 1), "no odd multiple of 2^-53"
    assert f[0] == 0.0 and f[-1] == 1.0 - 2.0 ** -53
    assert _draw(oracle, "fgo_standard_f64", 0x7FF) == 0.0
    assert _draw(oracle, "fgo_standard_f64", 0x800) == 2.0 ** -53
    rng = np.random.default_rng(3)
    for v in vs[::7]:
        w = (v & ~0x7FF) | int(rng.integers(0, 0x800))
        assert _draw(oracle, "fgo_standard_f64", w) == _draw(oracle, "fgo_standard_f64", v)


def test_open01_f64_bits(oracle):
    vs = _values()
    f = np.array([_draw(oracle, "fgo_open01_f64", v) for v in vs])
    assert np.all(np.diff(f) >= 0)
    k = f * 2.0 ** 53
    assert np.all(k == np.floor(k)) and np.all(k % 2 == 1), "not an odd multiple of 2^-53"
    assert f.min() == 2.0 ** -53 and f.max() == 1.0 - 2.0 ** -53
    assert _draw(oracle, "fgo_open01_f64", 0xFFF) == 2.0 ** -53
    assert _draw(oracle, "fgo_open01_f64", 0x1000) == 3.0 * 2.0 ** -53


@pytest.mark.parametrize("scale", [1.0, 0.1, 1.0 / 9.0])
def test_uniform_f32_sample_bits(oracle, scale):
    sc = float(np.float32(scale))
    vs = _values()
    f = np.array([_draw(oracle, "fgo_uniform_f32_sample", v, 0.0, sc) for v in vs], dtype=np.float32)
    assert np.all(np.diff(f) >= 0)
    assert f[0] == 0.0 and f[-1] < np.float32(sc)
    rng = np.random.default_rng(4)
    for v in vs[::5]:
        w = (v & ~((1 << 41) - 1)) | int(rng.integers(0, 1 << 41))
        assert _draw(oracle, "fgo_uniform_f32_sample", w, 0.0, sc) == _draw(oracle, "fgo_uniform_f32_sample",
                                                                             v, 0.0, sc)
    if scale == 1.0:
        k = f.astype(np.float64) * 2.0 ** 23
        assert np.all(k == np.floor(k))
        assert _draw(oracle, "fgo_uniform_f32_sample", (1 << 41) - 1, 0.0, 1.0) == 0.0
        assert _draw(oracle, "fgo_uniform_f32_sample", 1 << 41, 0.0, 1.0) == 2.0 ** -23
        assert f[-1] == 1.0 - 2.0 ** -23
    # low is added after the scaling
    assert _draw(oracle, "fgo_uniform_f32_sample", MASK, 5.0, sc) == np.float32(np.float32(f[-1]) + np.float32(5.0))


def test_pcg32_published_demo_vector(oracle):
    # pcg32_srandom_r(42, 54): inc = 54 * 2 + 1, state after seeding advances
    # from inc + 42; the demo prints the XSH-RR output of each state.  The
    # oracle's step advances first and outputs the new state, so starting it
    # at inc + 42 reproduces the demo's first round.
    want = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    inc = 54 * 2 + 1
    st = (C.c_uint64 * 1)(inc + 42)
    got = [oracle.lib().fgo_pcg32_next(st, inc) for _ in range(6)]
    assert got == want, [hex(g) for g in got]
