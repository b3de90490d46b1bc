"""Batched persistent renderer (fg_batch_*) on the GPU, through the C ABI.

Every plane of a batch must equal the single-plane device path bit for bit
(and hence the oracle), with and without graph capture, over repeated graph
replays interleaved with other work on the device; a run whose grains exceed
the planned capacity must render nothing and report FG_ECAPACITY.
"""
import numpy as np
import pytest

from film_grain_amd import capi, synth, workloads
from tests import fgcases

pytestmark = pytest.mark.gpu


def _cases(O):
    u = synth.sinusoid_field(41, 29, 21, n_waves=3)
    return [
        fgcases.c1_case(O, size=48, n_samples=64),
        fgcases.c2_case(O, w=40, h=24, n_samples=256),
        fgcases.build_case(O, "b_zoom1.5", u, 0.15, 0.15, "lognorm", zoom=1.5, n_samples=100, seed=99),
        fgcases.build_case(O, "b_n1", u[:7, :9], 0.2, n_samples=1, seed=3),
        fgcases.build_case(O, "b_tiny", np.full((1, 1), 0.6, np.float32), 0.1, n_samples=16, zoom=4.0, seed=6),
        fgcases.build_case(O, "b_n300_zoom0.5", u, 0.3, 0.05, "lognorm", n_samples=300, zoom=0.5, sigma=2.0, seed=10),
        fgcases.build_case(O, "b_rejection", np.full((6, 6), 0.99, np.float32), 0.05, n_samples=16, seed=12,
                           cell_delta=1.0),
    ]


def _tensors(torch, case):
    dev = torch.device("cuda:0")
    lam = torch.from_numpy(np.ascontiguousarray(case.lam)).to(dev)
    offs = torch.from_numpy(np.ascontiguousarray(case.offsets_input)).to(dev)
    return lam, offs


def _single(torch, ctx, case, lam, offs):
    out = torch.zeros((case.desc.out_h, case.desc.out_w), dtype=torch.float32, device="cuda:0")
    ctx.render_device("pixel", case.desc, lam, offs, out)
    ctx.stream_status()
    return out.cpu().numpy()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_planes_equal_single_plane_path(hip_ctx, oracle, graph):
    import torch
    cases = _cases(oracle)
    ins = [_tensors(torch, c) for c in cases]
    want = [_single(torch, hip_ctx, c, *t) for c, t in zip(cases, ins)]
    # the oracle agrees with the single-plane path on the small cases
    for c, w in zip(cases[:2], want[:2]):
        assert np.array_equal(w, oracle.render_pixelwise(c.plane, c.lam, c.offsets_input)), c.name
    outs = [torch.full((c.desc.out_h, c.desc.out_w), -1.0, dtype=torch.float32, device="cuda:0") for c in cases]
    batch = capi.Batch(hip_ctx, [(c.desc, t[0], t[1], o) for c, t, o in zip(cases, ins, outs)],
                       capture_graph=graph)
    info = batch.info()
    print(f"[batch graph={graph}] {info}", flush=True)
    assert info["graph"] == graph and info["items"] > 0
    for rep in range(3):
        for o in outs:
            o.fill_(-1.0)
        batch.run()
        batch.status()
        torch.cuda.synchronize()
        for c, o, w in zip(cases, outs, want):
            got = o.cpu().numpy()
            assert np.array_equal(got, w), f"replay {rep}: {c.name} max|d|={np.abs(got - w).max()}"
        # eager work on the device between replays
        _single(torch, hip_ctx, cases[0], *ins[0])
    batch.close()


def test_batch_capacity_guard(hip_ctx, oracle):
    import torch
    u = np.full((64, 64), 0.05, np.float32)
    case = fgcases.build_case(oracle, "cap_low", u, 0.1, n_samples=32, seed=5)
    lam, offs = _tensors(torch, case)
    out = torch.full((case.desc.out_h, case.desc.out_w), -1.0, dtype=torch.float32, device="cuda:0")
    batch = capi.Batch(hip_ctx, [(case.desc, lam, offs, out)], capture_graph=True)
    batch.run()
    batch.status()
    # far more grains than planned: the run renders nothing and says so
    hi = fgcases.build_case(oracle, "cap_high", np.full((64, 64), 0.95, np.float32), 0.1, n_samples=32, seed=5)
    lam.copy_(torch.from_numpy(np.ascontiguousarray(hi.lam)).to(lam.device))
    out.fill_(-1.0)
    batch.run()
    with pytest.raises(capi.FilmGrainError) as err:
        batch.status()
    assert err.value.code == capi.FG_ECAPACITY
    assert float(out.min()) == -1.0 and float(out.max()) == -1.0
    batch.close()
    # planned again on the new inputs, it renders them
    batch = capi.Batch(hip_ctx, [(case.desc, lam, offs, out)], capture_graph=True)
    batch.run()
    batch.status()
    want = _single(torch, hip_ctx, hi, lam, offs)
    assert np.array_equal(out.cpu().numpy(), want)
    batch.close()


def test_batch_uneven_images_equal_single_plane_path(hip_ctx):
    """A small version of the uneven batch (8 images, sides 64-320, mixed
    N, radius laws and channel counts) in one graph."""
    import torch
    images = workloads.batch_images(master_seed=7, n_images=8, min_side=64, max_side=320)
    planes = workloads.prepare_batch(images)
    dev = torch.device("cuda:0")
    ins = [(d, torch.from_numpy(l).to(dev), torch.from_numpy(o).to(dev)) for d, l, o in planes]
    want = []
    for d, lam, offs in ins:
        out = torch.zeros((d.out_h, d.out_w), dtype=torch.float32, device=dev)
        hip_ctx.render_device("pixel", d, lam, offs, out)
        want.append(out)
    hip_ctx.stream_status()
    outs = [torch.zeros_like(w) for w in want]
    batch = capi.Batch(hip_ctx, [(d, lam, offs, o) for (d, lam, offs), o in zip(ins, outs)], capture_graph=True)
    for _ in range(2):
        batch.run()
        batch.status()
        for i, (o, w) in enumerate(zip(outs, want)):
            assert torch.equal(o, w), f"plane {i}"
    batch.close()
