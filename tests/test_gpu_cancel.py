"""Cancellation (fg_cancel and the host-memory entry points' callbacks): a
raised flag stops the integrators at their next tile and the status reports
FG_ECANCELLED; lowered again, the same context renders the same planes as
before.  Every kernel still runs to its normal end: the flag is read once per
tile, never waited on."""
import time

import numpy as np
import pytest

from film_grain_amd import capi, workloads
from film_grain_amd.capi import FilmGrainError

pytestmark = pytest.mark.gpu


def _plane(algo, n=256, w=960, h=540, zoom=2.0):
    from film_grain_amd import host, synth
    u = synth.c2_rgb(7, w, h)[0]
    args = host.CliArgs(radius=0.1, radius_stddev=0.05, radius_dist="lognorm", zoom=zoom, iters=n, seed=7,
                        algo=algo)
    d, offs, offs_in = host.derive(args, w, h)
    nrm, _ = host.normalize_plane(u)
    lam = host.lambda_plane(nrm, d.inv_e_pi_r2)
    desc = capi.make_desc(d.in_w, d.in_h, d.out_w, d.out_h, d.n_samples, d.dist_kind, d.seed, d.zoom, d.delta,
                          d.rm, d.radius_mean, d.log_mu, d.log_sigma)
    return desc, lam, (offs_in if algo == "pixel" else offs)


@pytest.mark.parametrize("algo", ["pixel", "grain"])
def test_raised_flag_stops_device_render(hip_ctx, algo):
    import torch
    desc, lam, offs = _plane(algo)
    ld = torch.from_numpy(lam).to("cuda:0")
    od = torch.from_numpy(np.ascontiguousarray(offs)).to("cuda:0")
    ref = torch.empty((desc.out_h, desc.out_w), dtype=torch.float32, device="cuda:0")
    hip_ctx.render_device(algo, desc, ld, od, ref)
    hip_ctx.stream_status()
    out = torch.full_like(ref, -1.0)
    hip_ctx.cancel(True)
    try:
        hip_ctx.render_device(algo, desc, ld, od, out)
        with pytest.raises(FilmGrainError) as e:
            hip_ctx.stream_status()
        assert e.value.code == capi.FG_ECANCELLED
    finally:
        hip_ctx.cancel(False)
    # no tile ran: every output pixel keeps its fill
    assert bool((out == -1.0).all())
    hip_ctx.render_device(algo, desc, ld, od, out)
    hip_ctx.stream_status()
    assert torch.equal(out, ref)


def test_callback_cancels_a_host_render_in_flight(hip_ctx):
    desc, lam, offs = _plane("pixel", n=1024, w=1920, h=1080)
    t0 = time.perf_counter()
    ref = hip_ctx.render_host("pixel", desc, lam, offs)
    full = time.perf_counter() - t0
    calls = []

    def cancel():  # not on the check before launch; on every poll after it
        calls.append(1)
        return len(calls) > 1

    t0 = time.perf_counter()
    with pytest.raises(FilmGrainError) as e:
        hip_ctx.render_host("pixel", desc, lam, offs, cancel=cancel)
    cut = time.perf_counter() - t0
    print(f"[cancel] full render {full * 1e3:.1f} ms, cancelled after {cut * 1e3:.1f} ms, {len(calls)} polls",
          flush=True)
    assert e.value.code == capi.FG_ECANCELLED
    assert cut < full
    # the flag is lowered again: the context renders the same plane
    assert np.array_equal(hip_ctx.render_host("pixel", desc, lam, offs), ref)


def test_batch_run_cancelled_then_resumed(hip_ctx):
    import torch
    images = workloads.batch_images(master_seed=5, n_images=6, min_side=128, max_side=512)
    planes = workloads.prepare_batch(images)
    dev = "cuda:0"
    ins = [(d, torch.from_numpy(l).to(dev), torch.from_numpy(o).to(dev)) for d, l, o in planes]
    outs = [torch.full((d.out_h, d.out_w), -1.0, dtype=torch.float32, device=dev) for d, _, _ in ins]
    refs = []
    for (d, l, o), out in zip(ins, outs):
        r = torch.empty_like(out)
        hip_ctx.render_device("pixel", d, l, o, r)
        refs.append(r)
    hip_ctx.stream_status()
    batch = capi.Batch(hip_ctx, [(d, l, o, out) for (d, l, o), out in zip(ins, outs)], capture_graph=True)
    try:
        hip_ctx.cancel(True)
        try:
            batch.run()
            with pytest.raises(FilmGrainError) as e:
                batch.status()
            assert e.value.code == capi.FG_ECANCELLED
        finally:
            hip_ctx.cancel(False)
        assert all(bool((o == -1.0).all()) for o in outs)
        batch.run()
        batch.status()
        assert all(torch.equal(o, r) for o, r in zip(outs, refs))
    finally:
        batch.close()
