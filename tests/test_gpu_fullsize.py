"""The benchmark configurations at full size, on MI355X, against the oracle.

* C2 (1920x1080 -> 3840x2160, lognormal 0.1/0.05, N = 256): every cell of
  the plane's extended grid (2.1e8 cells, 4.5e7 grains) has the oracle's
  Poisson count, and its grains the oracle's bits (a per-cell digest of the
  f32 bits in draw order), src/pixelwise.rs:66-100.
* C3 (4096x4096 ramp, r = 0.05, N = 128, grain-wise): every input pixel's
  Poisson count (the rejection branch: lambda up to 880) and stream state
  after it equal the oracle's (src/grainwise.rs:37-45), and the whole frame
  equals the oracle's render (src/grainwise.rs:12-124) on all host cores.
* C4 (64 images of 128-2048 px, mixed radii, N and channels, seed 2024): the
  graph-captured batch's planes equal the single-plane path, and 16x16 tiles
  of eight planes equal the oracle.
* C5 (1920x1080 RGB, learned budgets 16-1024): the adaptive frame meets its
  RMSE target (fixed N = 1024 against the N = 4096 render), and sampled
  pixels equal the oracle's render at their own budgets.
"""
import os

import numpy as np
import pytest

from film_grain_amd import budget, capi, workloads
from tests import cellcheck, fgcases

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda:0"))


def _threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def test_c2_full_plane_every_cell(hip_ctx, oracle):
    case = fgcases.c2_case(oracle, w=1920, h=1080, n_samples=256, channel=0)
    print("[c2 cells] generating on the device", flush=True)
    rec, origin, size = hip_ctx.pixelwise_grains(case.desc, _dev(case.lam), _dev(case.offsets_input))
    print(f"[c2 cells] {len(rec)} grains over a {size[0]}x{size[1]} grid; digests", flush=True)
    q, d = cellcheck.device_cell_digests(rec, size)
    del rec
    print(f"[c2 cells] oracle on {_threads()} threads", flush=True)
    qo, do = oracle.cell_digests(case.plane, case.lam, origin, size, threads=_threads())
    bad_q = int(np.count_nonzero(q != qo))
    bad_d = int(np.count_nonzero(d != do))
    print(f"[c2 cells] {q.size} cells, {int(qo.sum())} grains: counts differ in {bad_q} cells, "
          f"grain bits in {bad_d}", flush=True)
    assert bad_q == 0 and bad_d == 0


def test_c3_full_plane_pixel_streams(hip_ctx, oracle):
    case = fgcases.c3_case(oracle, size=4096, n_samples=128)
    q, st = hip_ctx.grainwise_pixel_state(case.desc, _dev(case.lam))
    qo, sto = oracle.pixel_counts(case.plane, case.lam, threads=_threads(), states=True)
    bad_q = int(np.count_nonzero(q != qo))
    live = qo > 0
    bad_s = int(np.count_nonzero(np.any(st[live] != sto[live], axis=1)))
    print(f"[c3 pixels] {q.size} pixels, {int(qo.sum())} grains (lambda max {float(case.lam.max()):.1f}): "
          f"counts differ at {bad_q}, states at {bad_s}", flush=True)
    assert bad_q == 0 and bad_s == 0


def test_c3_full_frame_equals_oracle(hip_ctx, oracle):
    case = fgcases.c3_case(oracle, size=4096, n_samples=128)
    got = hip_ctx.render_host("grain", case.desc, case.lam, case.offsets)
    print(f"[c3 frame] device done; oracle on {_threads()} threads", flush=True)
    want = oracle.render_grainwise(case.plane, case.lam, case.offsets, threads=_threads())
    diff = np.abs(got.astype(np.float64) - want)
    print(f"[c3 frame] {got.size} pixels: {int(np.count_nonzero(diff))} differ, max {diff.max():.3g}", flush=True)
    assert np.array_equal(got, want)


def test_c4_full_batch(hip_ctx, oracle):
    import torch
    images = workloads.batch_images(master_seed=2024, n_images=64)
    planes = workloads.prepare_batch(images)
    dev = torch.device("cuda:0")
    ins = [(d, torch.from_numpy(l).to(dev), torch.from_numpy(o).to(dev)) for d, l, o in planes]
    px = sum(d.out_w * d.out_h * d.n_samples for d, _, _ in ins)
    print(f"[c4] {len(images)} images, {len(ins)} planes, {px:.4g} pixel-samples", flush=True)
    single = []
    for d, lam, offs in ins:
        out = torch.empty((d.out_h, d.out_w), dtype=torch.float32, device=dev)
        hip_ctx.render_device("pixel", d, lam, offs, out)
        single.append(out)
    hip_ctx.stream_status()
    outs = [torch.full_like(s, -1.0) for s in single]
    batch = capi.Batch(hip_ctx, [(d, lam, offs, o) for (d, lam, offs), o in zip(ins, outs)], capture_graph=True)
    print(f"[c4] batch {batch.info()}", flush=True)
    for rep in range(2):
        batch.run()
        batch.status()
        torch.cuda.synchronize()
        for i, (o, s) in enumerate(zip(outs, single)):
            assert torch.equal(o, s), f"replay {rep}: plane {i}"
    batch.close()
    rng = np.random.default_rng(4)
    picks = sorted(rng.choice(len(planes), size=8, replace=False).tolist())
    for i in picks:
        d, lam, offs = planes[i]
        p = fgcases.oracle_plane(oracle, d)
        for (x0, y0, x1, y1) in fgcases.tiles(rng, d.out_w, d.out_h, 16, 2):
            want = oracle.render_pixelwise(p, lam, offs, region=(x0, y0, x1, y1), threads=_threads())
            got = single[i][y0:y1, x0:x1].cpu().numpy()
            assert np.array_equal(got, want), f"plane {i} tile ({x0}, {y0}): max |d| {np.abs(got - want).max()}"
    print(f"[c4] planes {picks}: 16 oracle tiles equal", flush=True)


def test_c5_adaptive_frame(hip_ctx, oracle):
    import torch
    model = budget.ErrorModel.load()
    specs = budget.image_planes(50_000)
    planes = [budget.upload(s) for s in specs]
    refs = [budget.render_reference(hip_ctx, dp) for dp in planes]
    fixed = [budget.render_fixed(hip_ctx, dp, budget.C5["n_fixed"]) for dp in planes]
    target = budget._frame_rmse(fixed, refs)
    bufs = budget.frame_buffers(planes)
    outs, _ = budget.render_adaptive_frame(hip_ctx, model, planes, bufs, target ** 2, "learned")
    torch.cuda.synchronize()
    got = budget._frame_rmse(outs, refs)
    b = [(x.to(torch.int32) & 0xFFFF).reshape(dp.out_h, dp.out_w).cpu().numpy() for x, dp in zip(bufs.budget, planes)]
    mean_n = float(np.mean([x.mean() for x in b]))
    print(f"[c5] target {target * 255:.3f}/255, adaptive {got * 255:.3f}/255, mean budget {mean_n:.1f}", flush=True)
    assert got <= target * 1.05
    assert min(int(x.min()) for x in b) >= budget.C5["n_min"] and max(int(x.max()) for x in b) <= budget.C5["n_max"]
    # sampled pixels equal the oracle's render at their own budgets
    rng = np.random.default_rng(9)
    n_checked = 0
    for ci, dp in enumerate(planes):
        lam = dp.lam.cpu().numpy()
        offs = dp.offs.cpu().numpy()
        out = outs[ci].cpu().numpy()
        ys = rng.integers(0, dp.out_h, 24)
        xs = rng.integers(0, dp.out_w, 24)
        for y, x in zip(ys, xs):
            n = int(b[ci][y, x])
            p = fgcases.oracle_plane(oracle, dp.desc, n_samples=n)
            want = oracle.render_pixelwise(p, lam, np.ascontiguousarray(offs[:n]), region=(x, y, x + 1, y + 1),
                                           threads=1)
            assert out[y, x] == want[0, 0], f"plane {ci} pixel ({x}, {y}) budget {n}"
            n_checked += 1
    print(f"[c5] {n_checked} sampled pixels equal the oracle at their budgets", flush=True)
