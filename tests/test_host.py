"""CPU tests of the product's host layer (C++ host library through its C ABI)
and of the C ABI surface itself; no GPU needed.

The host restates src/params.rs, src/model.rs, src/rng.rs, src/choose.rs and
src/color.rs; its values are checked against the oracle, which is pinned by
tests/test_oracle_kats.py, and its error messages against the reference's
strings (src/params.rs:263-532).
"""
import re
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

from film_grain_amd import capi, host, synth

ROOT = Path(__file__).resolve().parents[1]


# ---------------------------------------------------------------- ABI surface
def _declared(header: Path):
    txt = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    names = set(re.findall(r"\b(fgh?_[a-z0-9_]+)\s*\(", txt))
    return {n for n in names if not n.endswith("_fn")}


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    declared = _declared(ROOT / "include/film_grain_hip.h") | _declared(ROOT / "include/film_grain.h")
    assert declared, "no declarations parsed"
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert set(capi.EXPORTED) <= declared and set(host.EXPORTED) <= declared


def test_version_string():
    assert b"gfx950" in capi.lib().fg_version()


def test_context_without_device_reports_enodev():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.FilmGrainError) as e:
        capi.Context(0)
    assert e.value.code == capi.FG_ENODEV


# ---------------------------------------------------------------- host vs oracle
CONFIGS = [
    dict(radius=0.1, radius_stddev=0.05, radius_dist="lognorm", zoom=2.0, iters=256, seed=7),
    dict(radius=0.12, zoom=1.0, iters=64, seed=42),
    dict(radius=0.05, zoom=1.0, iters=128, seed=11),
    dict(radius=0.3, radius_stddev=0.15, radius_dist="lognorm", zoom=1.5, iters=33, seed=123456789),
    dict(radius=0.25, radius_stddev=0.0, radius_dist="lognorm", zoom=0.7, iters=5, seed=1,
         max_radius="quantile=0.9", cell="0.07"),
    dict(radius=0.2, zoom=3.0, iters=17, seed=3, max_radius="0.5", cell="auto", sigma=2.5),
]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_derive_and_offsets_match_oracle(oracle, cfg):
    a = host.CliArgs(**cfg)
    d, offs, offs_in = host.derive(a, 333, 211)
    mr = a.max_radius.replace("=", " ").split()
    mrs = ("quantile", float(mr[1]) if len(mr) > 1 else 0.999) if mr[0] == "quantile" else ("absolute", float(mr[0]))
    cell = None if a.cell in ("approx mu_r", "auto", "mu_r") else float(a.cell)
    od = oracle.derive(a.radius, a.radius_stddev, a.radius_dist, a.zoom, a.sigma, a.iters,
                       max_radius=mrs, cell_delta=cell, seed=a.seed, in_w=333, in_h=211)
    for f in ("rm", "delta", "inv_e_pi_r2", "log_mu", "log_sigma", "out_w", "out_h"):
        assert getattr(d, f) == getattr(od, f), f
    oo = oracle.make_offsets(a.seed, a.iters, a.sigma)
    assert np.array_equal(offs, oo)
    assert np.array_equal(offs_in, (oo / np.float32(a.zoom)).astype(np.float32))
    assert d.algorithm == oracle.lib().fgo_choose_algorithm(a.radius, d.radius_stddev, d.rm, a.iters, 0)


def test_offsets_edge_cases_match_oracle(oracle):
    for seed, n, sigma in [(0, 1, 0.8), (2**64 - 1, 1000, 0.8), (5, 3, 1e-9), (9, 257, 100.0)]:
        assert np.array_equal(host.make_offsets(seed, n, sigma), oracle.make_offsets(seed, n, sigma))


def test_lambda_prepass_matches_oracle(oracle):
    rng = np.random.default_rng(0)
    u = rng.uniform(-0.2, 1.5, size=4096).astype(np.float32)
    u[::97] = np.nan
    for field in (u, synth.c2_rgb(7, 64, 48)[2].ravel(), np.clip(u, 0, 0.999)):
        a, ma = host.normalize_plane(field)
        b, mb = oracle.normalize_plane(field)
        assert np.array_equal(a, b, equal_nan=True) and (ma == mb)
        for inv in (25.46479, 127.32395, 2.829):
            assert np.array_equal(host.lambda_plane(a, inv), oracle.lambda_plane(b, inv), equal_nan=True)


def test_two_inverse_normal_implementations_agree(oracle):
    for p in np.concatenate([np.linspace(1e-12, 1 - 1e-12, 101), [0.999, 0.975, 0.5, 1e-300]]):
        a, b = host.inverse_normal_cdf(p), oracle.lib().fgo_inverse_normal_cdf(p)
        assert a == pytest.approx(b, rel=4e-15, abs=1e-15)


# ---------------------------------------------------------------- validation
ERRORS = [
    (dict(radius=0.0), "radius: value must be greater than 0"),
    (dict(radius=float("inf")), "radius: value must be finite"),
    (dict(radius_stddev=-0.1), "radius-stddev: value must be >= 0"),
    (dict(zoom=0.0), "zoom: value must be greater than 0"),
    (dict(sigma=-1.0), "sigma: value must be greater than 0"),
    (dict(radius=1e39), "radius: value is too large for single precision"),
    (dict(cell="abc"), "cell: expected a positive number or 'approx mu_r'"),
    (dict(cell="-1"), "cell: cell size must be > 0"),
    (dict(max_radius="quantile 1.5"), "max-radius: quantile must lie in the open interval (0,1)"),
    (dict(max_radius="quantile: x"), "max-radius: expected numeric quantile value in (0,1)"),
    (dict(max_radius="abc"), "max-radius: expected a positive number or 'quantile'"),
    (dict(max_radius="-2"), "max-radius: absolute radius must be > 0"),
    (dict(max_radius="nan"), "max-radius: absolute radius must be finite"),
    (dict(roi="1,2"), "roi: expected X0,Y0:X1,Y1"),
    (dict(roi="1,2:3,4:5"), "roi: unexpected extra ':'"),
    (dict(roi="5,5:1,1"), "roi: roi end must be greater than start (exclusive bounds)"),
    (dict(roi="1:2"), "roi: expected two comma-separated integers"),
    (dict(roi="1,2,3:4,5"), "roi: too many commas"),
    (dict(roi="a,b:1,2"), "roi: expected an unsigned integer without suffixes"),
    (dict(size=""), "size: size string cannot be empty"),
    (dict(size="x480"), "size: height-only form is not supported; provide width or WxH"),
    (dict(size="x"), "size: missing width and height around separator"),
    (dict(size="0"), "size: output width must be > 0"),
    (dict(size="640x0"), "size: output height must be > 0"),
    (dict(size="12k"), "size: expected an unsigned integer without suffixes"),
]


@pytest.mark.parametrize("kw,msg", ERRORS)
def test_parameter_errors_match_reference_strings(kw, msg):
    with pytest.raises(host.RenderError) as e:
        host.derive(host.CliArgs(**kw), 10, 10)
    assert e.value.code == host.FGH_EPARAMS
    assert str(e.value) == f"parameter error: {msg}"


def test_size_and_roi_forms():
    d, _, _ = host.derive(host.CliArgs(size="640"), 320, 200)
    assert (d.out_w, d.out_h) == (640, 400)
    d, _, _ = host.derive(host.CliArgs(size=" 640 X 50 "), 320, 200)
    assert (d.out_w, d.out_h) == (640, 50)
    d, _, _ = host.derive(host.CliArgs(size="640x"), 3, 2)
    assert (d.out_w, d.out_h) == (640, 427)
    d, _, _ = host.derive(host.CliArgs(zoom=1.5), 7, 5)
    assert (d.out_w, d.out_h) == (11, 8)


def test_empty_input_is_a_message_error():
    with pytest.raises(host.RenderError) as e:
        host.derive(host.CliArgs(), 0, 10)
    assert e.value.code == host.FGH_EMESSAGE and "empty" in str(e.value)


def test_cpu_device_is_unsupported_and_no_gpu_is_reported():
    rgb = np.full((8, 8, 3), 0.5, np.float32)
    with pytest.raises(host.RenderError) as e:
        host.render_rgb(host.CliArgs(device="cpu"), rgb)
    assert e.value.code == host.FGH_EUNSUPPORTED
    import torch
    if not torch.cuda.is_available():
        with pytest.raises(host.RenderError) as e:
            host.render_rgb(host.CliArgs(device="hip"), rgb)
        assert e.value.code == host.FGH_EGPU and "no compatible GPU adapter found" in str(e.value)


def test_cancel_before_work_is_reported():
    rgb = np.full((8, 8, 3), 0.5, np.float32)
    with pytest.raises(host.RenderError) as e:
        host.render_rgb(host.CliArgs(device="hip"), rgb, cancel=lambda: True)
    assert e.value.code == host.FGH_ECANCELLED


# ---------------------------------------------------------------- colour and files
def test_rgb_roundtrip_is_exact_on_u8_levels():
    v = np.arange(256, dtype=np.float32) / np.float32(255.0)
    rgb = np.stack([v, v[::-1], np.roll(v, 77)], axis=-1).reshape(16, 16, 3)
    out = host.color_roundtrip(rgb, "rgb")
    assert np.array_equal(out, np.rint(rgb * 255).astype(np.uint8))


def test_luma_roundtrip_recovers_colour_within_one_level():
    rng = np.random.default_rng(3)
    rgb = (rng.integers(0, 256, size=(32, 32, 3)) / 255.0).astype(np.float32)
    out = host.color_roundtrip(rgb, "luma").astype(int)
    assert np.abs(out - np.rint(rgb * 255)).max() <= 1


def test_host_colour_steps_equal_oracle(oracle):
    """The host restatement's split and split-then-merge against the
    oracle's (pinned by BT.709 closed forms in test_oracle_colour)."""
    rng = np.random.default_rng(8)
    rgb = rng.uniform(-0.1, 1.1, (23, 31, 3)).astype(np.float32)
    rgb[2, 3] = np.nan
    for name, mode in (("luma", oracle.COLOR_LUMA), ("rgb", oracle.COLOR_RGB)):
        want = oracle.color_split(rgb, mode)
        assert np.array_equal(host.color_split(rgb, name), want, equal_nan=True), name
        assert np.array_equal(host.color_roundtrip(rgb, name), oracle.color_merge(mode, *want)), name


def _png(path, arr, bitdepth=8, ctype=0):
    h, w = arr.shape[:2]
    dt = ">u2" if bitdepth == 16 else "u1"
    raw = b"".join(b"\x00" + np.ascontiguousarray(row).astype(dt).tobytes() for row in arr)
    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    data = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, bitdepth, ctype, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw, 3)) + chunk(b"IEND", b""))
    Path(path).write_bytes(data)


def test_image_files_decode_like_to_rgb32f(tmp_path):
    rng = np.random.default_rng(1)
    g8 = rng.integers(0, 256, size=(9, 13), dtype=np.uint8)
    _png(tmp_path / "g8.png", g8)
    img = host.load_image(str(tmp_path / "g8.png"))
    assert np.array_equal(img, np.repeat((g8 / np.float32(255))[..., None], 3, axis=2).astype(np.float32))
    c16 = rng.integers(0, 65536, size=(5, 7, 3)).astype(np.uint16)
    _png(tmp_path / "c16.png", c16, bitdepth=16, ctype=2)
    img = host.load_image(str(tmp_path / "c16.png"))
    assert np.allclose(img, c16 / np.float32(65535), rtol=0, atol=0)
    ppm = b"P6\n7 5\n255\n" + c16.astype(np.uint8).tobytes()
    (tmp_path / "c.ppm").write_bytes(ppm)
    assert np.array_equal(host.load_image(str(tmp_path / "c.ppm")), c16.astype(np.uint8) / np.float32(255))
    f = rng.uniform(0, 1, size=(4, 6, 3)).astype("<f4")
    (tmp_path / "f.pfm").write_bytes(b"PF\n6 4\n-1.0\n" + f[::-1].tobytes())
    assert np.array_equal(host.load_image(str(tmp_path / "f.pfm")), f)


def test_png_writer_roundtrip(tmp_path):
    rng = np.random.default_rng(2)
    rgb8 = rng.integers(0, 256, size=(11, 17, 3), dtype=np.uint8)
    host.save_rgb8(str(tmp_path / "o.png"), rgb8, "png")
    assert np.array_equal(host.load_image(str(tmp_path / "o.png")), rgb8 / np.float32(255))
    host.save_rgb8(str(tmp_path / "o.ppm"), rgb8, "ppm")
    assert np.array_equal(host.load_image(str(tmp_path / "o.ppm")), rgb8 / np.float32(255))
    with pytest.raises(host.RenderError) as e:
        host.save_rgb8(str(tmp_path / "o.jpg"), rgb8, "jpg")
    assert e.value.code == host.FGH_EUNSUPPORTED


def test_bad_image_is_an_image_error(tmp_path):
    (tmp_path / "x.png").write_bytes(b"not an image")
    with pytest.raises(host.RenderError) as e:
        host.load_image(str(tmp_path / "x.png"))
    assert e.value.code == host.FGH_EIMAGE
    with pytest.raises(host.RenderError) as e:
        host.load_image(str(tmp_path / "missing.png"))
    assert e.value.code == host.FGH_EIO
