"""Frames on the device (fg_render_frame*, fg_lambda_device, fg_color_*_device,
fg_log_f32_device) against the host path, bit for bit: the f32 logarithm over
its whole domain, the lambda pre-pass, the colour split and merge, and whole
frames (luma with the chroma resize, RGB, both algorithms, an ROI); and the
colour steps and whole frames against the oracle's restatement of
src/color.rs, which is pinned independently of the host code."""
import numpy as np
import pytest

from film_grain_amd import capi, host
from oracle import oracle as O
from tests import fgcases

pytestmark = pytest.mark.gpu


def test_log_f32_matches_the_c_library_on_its_whole_domain(hip_ctx):
    import torch
    lo = np.array([1e-6], np.float32).view(np.uint32)[0]
    hi = np.array([1.0], np.float32).view(np.uint32)[0]
    bad = 0
    # every float in [1e-6, 1], in four chunks
    edges = np.linspace(int(lo), int(hi) + 1, 5).astype(np.int64)
    for a, b in zip(edges[:-1], edges[1:]):
        x = np.arange(a, b, dtype=np.int64).astype(np.uint32).view(np.float32)
        want = host.logf_array(x)
        xd = torch.from_numpy(x).to("cuda:0")
        got = torch.empty_like(xd)
        hip_ctx.log_f32_device(xd, got)
        g = got.cpu().numpy()
        bad += int(np.sum(g.view(np.uint32) != want.view(np.uint32)))
        del xd, got
    print(f"[frame] logf: {int(hi) - int(lo) + 1} inputs, {bad} differ", flush=True)
    assert bad == 0


def _planes(rng, h, w):
    u = rng.random((h, w), dtype=np.float32)
    yield "unit", u
    yield "scaled", (u * 3.7).astype(np.float32)          # normalised by the maximum
    v = u.copy()
    v[::7, ::5] = np.nan
    v[1::9, ::3] = -0.25
    v[2, :] = 0.0
    v[3, :] = np.float32(1.0 - 1e-6)
    v[4, :] = 1.0
    yield "special", v
    yield "all_negative", -u
    yield "tiny", (u * 1e-7).astype(np.float32)


def test_lambda_pre_pass_equals_host(hip_ctx):
    import torch
    rng = np.random.default_rng(1)
    for inv in (3.18, 31.8, 0.0):
        for name, u in _planes(rng, 37, 53):
            nrm, _ = host.normalize_plane(u)
            lam = host.lambda_plane(nrm, inv)
            ud = torch.from_numpy(np.ascontiguousarray(u)).to("cuda:0")
            nd = torch.empty_like(ud)
            ld = torch.empty_like(ud)
            hip_ctx.lambda_device(ud, inv, ld, norm=nd)
            torch.cuda.synchronize()
            assert np.array_equal(nd.cpu().numpy(), nrm, equal_nan=True), (name, inv)
            assert np.array_equal(ld.cpu().numpy().view(np.uint32), lam.view(np.uint32)), (name, inv)


def test_colour_split_and_merge_equal_host(hip_ctx):
    import torch
    rng = np.random.default_rng(2)
    rgb = rng.uniform(-0.1, 1.1, (29, 41, 3)).astype(np.float32)
    rgb[0, 0] = np.nan
    for mode, name in ((capi.COLOR_LUMA, "luma"), (capi.COLOR_RGB, "rgb")):
        want = host.color_split(rgb, name)
        rd = torch.from_numpy(rgb).to("cuda:0")
        p = [torch.empty((29, 41), dtype=torch.float32, device="cuda:0") for _ in range(3)]
        hip_ctx.color_split_device(mode, rd, *p)
        out8 = torch.empty((29, 41, 3), dtype=torch.uint8, device="cuda:0")
        hip_ctx.color_merge_device(mode, p[0], p[1], p[2], out8)
        torch.cuda.synchronize()
        for i in range(3):
            assert np.array_equal(p[i].cpu().numpy(), want[i], equal_nan=True), (name, i)
        assert np.array_equal(out8.cpu().numpy(), host.color_roundtrip(rgb, name)), name


FRAMES = [
    dict(color_mode="luma", zoom=2.0, radius=0.1, radius_stddev=0.05, radius_dist="lognorm", iters=64,
         algo="pixel"),
    dict(color_mode="rgb", zoom=1.0, radius=0.12, iters=32, algo="pixel"),
    dict(color_mode="rgb", zoom=1.5, radius=0.05, iters=16, algo="grain"),
    dict(color_mode="luma", zoom=0.5, radius=0.2, iters=24, algo="pixel", roi="3,2:45,30"),
    dict(color_mode="luma", zoom=1.0, radius=0.05, iters=16, algo="grain", seed=99),
]


@pytest.mark.parametrize("i", range(len(FRAMES)))
def test_frame_on_device_equals_host_path(hip_ctx, i):
    import torch
    kw = dict(FRAMES[i])
    args = host.CliArgs(**kw)
    rng = np.random.default_rng(10 + i)
    rgb = rng.random((34, 48, 3), dtype=np.float32)
    rgb[5:12, 7:20] = 0.0
    rgb[20:25, 30:40] = 0.98
    a8, ap, _ = host.render_rgb(args, rgb)
    b8, bp, _ = host.render_rgb(args, rgb, frame_on_device=True)
    print(f"[frame] {kw}: out {a8.shape}, planes {ap.shape}", flush=True)
    assert np.array_equal(ap, bp), kw
    assert np.array_equal(a8, b8), kw
    if "roi" in kw:
        return
    # the device-resident entry point, from device buffers
    d, offs, offs_in = host.derive(args, 48, 34)
    fd = capi.FrameDesc()
    fd.plane = capi.make_desc(d.in_w, d.in_h, d.out_w, d.out_h, d.n_samples, d.dist_kind, d.seed, d.zoom, d.delta,
                              d.rm, d.radius_mean, d.log_mu, d.log_sigma)
    fd.color_mode = capi.COLOR_RGB if kw["color_mode"] == "rgb" else capi.COLOR_LUMA
    fd.algo = capi.ALGO_PIXEL if kw["algo"] == "pixel" else capi.ALGO_GRAIN
    fd.inv_e_pi_r2 = d.inv_e_pi_r2
    rd = torch.from_numpy(rgb).to("cuda:0")
    od = torch.from_numpy(np.ascontiguousarray(offs_in if kw["algo"] == "pixel" else offs)).to("cuda:0")
    c8 = torch.empty((d.out_h, d.out_w, 3), dtype=torch.uint8, device="cuda:0")
    cp = torch.empty(ap.shape, dtype=torch.float32, device="cuda:0")
    hip_ctx.render_frame_device(fd, rd, od, c8, cp)
    hip_ctx.stream_status()
    assert np.array_equal(cp.cpu().numpy(), ap)
    assert np.array_equal(c8.cpu().numpy(), a8)


def test_colour_split_and_merge_equal_oracle(hip_ctx):
    """Device split and merge against the oracle: NaN and out-of-range
    inputs, and the luma merge with Cb/Cr resized nearest from other sizes."""
    import torch
    rng = np.random.default_rng(12)
    rgb = rng.uniform(-0.1, 1.1, (37, 53, 3)).astype(np.float32)
    rgb[0, 0] = np.nan
    rgb[1, 1, 2] = np.inf
    rd = torch.from_numpy(rgb).to("cuda:0")
    for mode in (O.COLOR_LUMA, O.COLOR_RGB):
        want = O.color_split(rgb, mode)
        p = [torch.empty((37, 53), dtype=torch.float32, device="cuda:0") for _ in range(3)]
        hip_ctx.color_split_device(mode, rd, *p)
        torch.cuda.synchronize()
        for i in range(3):
            assert np.array_equal(p[i].cpu().numpy(), want[i], equal_nan=True), (mode, i)
    y, cb, cr = O.color_split(rgb, O.COLOR_LUMA)
    for (h, w) in ((74, 106), (55, 79), (18, 26), (37, 53), (41, 50)):
        yy = rng.uniform(-0.05, 1.05, (h, w)).astype(np.float32)
        want = O.color_merge(O.COLOR_LUMA, yy, cb, cr)
        out8 = torch.empty((h, w, 3), dtype=torch.uint8, device="cuda:0")
        hip_ctx.color_merge_device(capi.COLOR_LUMA, torch.from_numpy(yy).to("cuda:0"),
                                   torch.from_numpy(cb).to("cuda:0"), torch.from_numpy(cr).to("cuda:0"), out8)
        torch.cuda.synchronize()
        assert np.array_equal(out8.cpu().numpy(), want), (h, w)
    planes = rng.uniform(-0.05, 1.05, (3, 37, 53)).astype(np.float32)
    planes[1, 2, 3] = np.nan
    out8 = torch.empty((37, 53, 3), dtype=torch.uint8, device="cuda:0")
    pd = [torch.from_numpy(np.ascontiguousarray(planes[i])).to("cuda:0") for i in range(3)]
    hip_ctx.color_merge_device(capi.COLOR_RGB, *pd, out8)
    torch.cuda.synchronize()
    assert np.array_equal(out8.cpu().numpy(), O.color_merge(O.COLOR_RGB, *planes))


@pytest.mark.parametrize("i", [i for i in range(len(FRAMES)) if "roi" not in FRAMES[i]])
def test_frame_on_device_equals_oracle(hip_ctx, i):
    """A whole frame through the device entry point against a frame the
    oracle alone composes (split, lambda, integrator, merge)."""
    kw = dict(FRAMES[i])
    args = host.CliArgs(**kw)
    rng = np.random.default_rng(10 + i)
    rgb = rng.random((34, 48, 3), dtype=np.float32)
    rgb[5:12, 7:20] = 0.0
    rgb[20:25, 30:40] = 0.98
    got8, got_planes, _ = host.render_rgb(args, rgb, frame_on_device=True)
    want8, want_planes = fgcases.oracle_frame(O, rgb, **{k: v for k, v in kw.items()})
    print(f"[frame-oracle] {kw}: out {got8.shape}", flush=True)
    assert np.array_equal(got_planes, want_planes), kw
    assert np.array_equal(got8, want8), kw
