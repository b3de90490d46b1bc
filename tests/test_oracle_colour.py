"""Pins of the oracle's colour restatement (src/color.rs, src/model.rs:77-98)
by what BT.709 and the rounding rules fix, independently of the formulas'
text: the constants' defining identities, closed forms on pure colours, the
exact inverse over every 8-bit colour, to_u8 on every level, and nearest
resize as pixel replication.  A swapped constant, a wrong denominator or a
round-half-even resize fails one of them."""
import numpy as np
import pytest

from oracle import oracle as O


def test_constants_are_bt709():
    kr, kg, kb, cb_d, cr_d = (float(v) for v in O.color_constants())
    # Rec. ITU-R BT.709-6 item 3.2: E'Y = 0.2126 R + 0.7152 G + 0.0722 B
    assert (kr, kg, kb) == (float(np.float32(0.2126)), float(np.float32(0.7152)), float(np.float32(0.0722)))
    assert round(0.2126 + 0.7152 + 0.0722, 12) == 1.0
    # Cb = (B - Y) / (2 (1 - Kb)), Cr = (R - Y) / (2 (1 - Kr))
    assert cb_d == float(np.float32(2.0 * (1.0 - 0.0722)))
    assert cr_d == float(np.float32(2.0 * (1.0 - 0.2126)))


def test_pure_colours_split_to_their_closed_forms():
    rgb = np.array([[[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [0, 0, 0], [0.5, 0.5, 0.5]]], np.float32)
    y, cb, cr = O.color_split(rgb, O.COLOR_LUMA)[:, 0, :]
    assert y[0] == np.float32(0.2126) and y[1] == np.float32(0.7152) and y[2] == np.float32(0.0722)
    assert abs(y[3] - 1.0) <= 2 ** -23 and y[3] <= 1.0 and y[4] == 0.0 and abs(y[5] - 0.5) <= 2 ** -23
    # the extremes of the chroma axes: blue has Cb = +1/2, red has Cr = +1/2
    assert abs(cb[2] - 0.5) <= 2 ** -24 and abs(cr[0] - 0.5) <= 2 ** -24
    # their opposites (yellow: Cb = -1/2, cyan: Cr = -1/2) through the same law
    yc = O.color_split(np.array([[[1, 1, 0], [0, 1, 1]]], np.float32), O.COLOR_LUMA)[:, 0, :]
    assert abs(yc[1, 0] + 0.5) <= 2 ** -23 and abs(yc[2, 1] + 0.5) <= 2 ** -23
    assert abs(cb[0] + 0.2126 / 1.8556) <= 1e-7 and abs(cr[2] + 0.0722 / 1.5748) <= 1e-7
    # greys carry no chroma
    assert max(abs(cb[3]), abs(cr[3]), abs(cb[5]), abs(cr[5])) <= 1e-7 and cb[4] == 0.0 and cr[4] == 0.0


def test_rgb_mode_clamps_and_keeps_channels():
    rgb = np.array([[[-0.5, 0.25, 1.5], [np.nan, 2.0, -0.0]]], np.float32)
    r, g, b = O.color_split(rgb, O.COLOR_RGB)[:, 0, :]
    assert r[0] == 0.0 and g[0] == 0.25 and b[0] == 1.0
    assert np.isnan(r[1]) and g[1] == 1.0 and b[1] == 0.0  # f32::clamp keeps NaN


def test_split_then_merge_returns_every_8bit_colour():
    k = np.arange(1 << 24, dtype=np.uint32)
    rgb8 = np.stack([(k >> 16) & 255, (k >> 8) & 255, k & 255], -1).astype(np.uint8).reshape(4096, 4096, 3)
    rgb = (rgb8 / np.float32(255)).astype(np.float32)
    for mode in (O.COLOR_LUMA, O.COLOR_RGB):
        p = O.color_split(rgb, mode)
        assert np.array_equal(O.color_merge(mode, p[0], p[1], p[2]), rgb8), mode


def test_to_u8_levels_and_saturation():
    levels = np.arange(256, dtype=np.float32) / np.float32(255)
    assert np.array_equal(O.to_u8(levels), np.arange(256, dtype=np.uint8))
    # floor(255 v + 1/2): rounds half up between levels
    k = np.arange(255, dtype=np.float32)
    assert np.array_equal(O.to_u8((k + np.float32(0.49)) / np.float32(255)), k.astype(np.uint8))
    assert np.array_equal(O.to_u8((k + np.float32(0.51)) / np.float32(255)), (k + 1).astype(np.uint8))
    assert list(O.to_u8([-1.0, -0.0, 1.0, 7.0, np.inf, -np.inf, np.nan])) == [0, 0, 255, 255, 255, 0, 0]


@pytest.mark.parametrize("f", [2, 3, 4])
def test_resize_nearest_replicates_at_integer_zoom(f):
    rng = np.random.default_rng(f)
    src = rng.random((7, 11), dtype=np.float32)
    out = O.resize_nearest(src, 11 * f, 7 * f)
    assert np.array_equal(out, np.repeat(np.repeat(src, f, axis=0), f, axis=1))


def test_resize_nearest_identity_halving_and_empty():
    rng = np.random.default_rng(9)
    src = rng.random((8, 10), dtype=np.float32)
    assert np.array_equal(O.resize_nearest(src, 10, 8), src)
    # (x + 0.5) * 2 - 0.5 = 2x + 0.5 rounds half away from zero (f32::round)
    assert np.array_equal(O.resize_nearest(src, 5, 4), src[1::2, 1::2])
    assert O.resize_nearest(src, 0, 3).shape == (3, 0)
    assert np.array_equal(O.resize_nearest(np.zeros((0, 4), np.float32), 3, 2), np.zeros((2, 3), np.float32))


def test_luma_merge_resizes_chroma():
    rng = np.random.default_rng(4)
    rgb = rng.random((6, 5, 3), dtype=np.float32)
    y, cb, cr = O.color_split(rgb, O.COLOR_LUMA)
    y2 = np.repeat(np.repeat(y, 2, axis=0), 2, axis=1)
    want = np.repeat(np.repeat(O.color_merge(O.COLOR_LUMA, y, cb, cr), 2, axis=0), 2, axis=1)
    assert np.array_equal(O.color_merge(O.COLOR_LUMA, y2, cb, cr), want)
