"""The planes of one frame rendered together (fg_render_pixelwise_planes_device),
which hashes each cell's stream once for all planes: every plane equals the
single-plane render of the same description bit for bit, whatever branch of
the Poisson sampler each plane's cells take, and random tiles equal the
oracle."""
import numpy as np
import pytest

from film_grain_amd import synth
from tests import fgcases

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _render_both(torch, ctx, cases):
    desc = cases[0].desc
    offs = _dev(torch, cases[0].offsets_input)
    lams = [_dev(torch, c.lam) for c in cases]
    single = []
    for lam in lams:
        out = torch.full((desc.out_h, desc.out_w), -1.0, dtype=torch.float32, device="cuda:0")
        ctx.render_device("pixel", desc, lam, offs, out)
        single.append(out)
    joint = [torch.full((desc.out_h, desc.out_w), -2.0, dtype=torch.float32, device="cuda:0") for _ in cases]
    ctx.render_planes_device(desc, lams, offs, joint)
    ctx.stream_status()
    return [s.cpu().numpy() for s in single], [j.cpu().numpy() for j in joint]


def _planes(O, fields, **kw):
    return [fgcases.build_case(O, f"p{i}", u, **kw) for i, u in enumerate(fields)]


@pytest.mark.parametrize("n", [1, 2, 3])
def test_joint_planes_equal_single_planes(hip_ctx, oracle, n):
    import torch
    rgb = synth.c2_rgb(7, 120, 80)
    cases = _planes(oracle, rgb[:n], radius=0.1, stddev=0.05, dist="lognorm", zoom=2.0, n_samples=64, seed=7)
    single, joint = _render_both(torch, hip_ctx, cases)
    rng = np.random.default_rng(n)
    for c, a, b in zip(cases, single, joint):
        assert np.array_equal(a, b), c.name
        for (x0, y0, x1, y1) in fgcases.tiles(rng, c.derived.out_w, c.derived.out_h, 16, 3):
            want = oracle.render_pixelwise(c.plane, c.lam, c.offsets_input, region=(x0, y0, x1, y1))
            assert np.array_equal(a[y0:y1, x0:x1], want), (c.name, x0, y0)


def test_joint_planes_mixed_branches(hip_ctx, oracle):
    """One plane dark (every cell skipped), one bright enough with a large
    cell for the rejection branch, and a ramp whose cells take both the Knuth
    and the rejection branch."""
    import torch
    h, w = 40, 56
    ramp = np.tile(np.linspace(0.0, 0.999, w, dtype=np.float32), (h, 1))
    dark = np.zeros((h, w), np.float32)
    bright = np.full((h, w), 0.995, np.float32)
    cases = _planes(oracle, [dark, bright, ramp], radius=0.2, n_samples=32, seed=12, cell_delta=0.9)
    single, joint = _render_both(torch, hip_ctx, cases)
    for c, a, b in zip(cases, single, joint):
        assert np.array_equal(a, b), c.name
    assert np.array_equal(single[1], oracle.render_pixelwise(cases[1].plane, cases[1].lam, cases[1].offsets_input))
    assert not np.any(single[0])


def test_joint_planes_full_c2_frame(hip_ctx, oracle):
    import torch
    cases = [fgcases.c2_case(oracle, w=1920, h=1080, n_samples=256, channel=c) for c in range(3)]
    single, joint = _render_both(torch, hip_ctx, cases)
    for c, a, b in zip(cases, single, joint):
        print(f"[planes] {c.name}: mean {a.mean():.5f}", flush=True)
        assert np.array_equal(a, b), c.name


def test_batch_groups_an_images_planes(hip_ctx, oracle):
    """A batch whose consecutive planes share a description and offsets (one
    RGB image), between planes of other descriptions: the grouped planes are
    counted in one pass and every plane still equals its single-plane render."""
    import torch
    from film_grain_amd import capi
    rgb = synth.c2_rgb(9, 72, 48)
    image = _planes(oracle, rgb, radius=0.1, stddev=0.05, dist="lognorm", zoom=2.0, n_samples=64, seed=9)
    other = [fgcases.c1_case(oracle, size=40, n_samples=32), fgcases.build_case(oracle, "solo", rgb[0], 0.15,
                                                                                 n_samples=48, seed=3)]
    cases = [other[0]] + image + [other[1]]
    ins = [(_dev(torch, c.lam), _dev(torch, c.offsets_input)) for c in cases]
    want = []
    for c, (lam, offs) in zip(cases, ins):
        out = torch.zeros((c.desc.out_h, c.desc.out_w), dtype=torch.float32, device="cuda:0")
        hip_ctx.render_device("pixel", c.desc, lam, offs, out)
        want.append(out)
    outs = [torch.full((c.desc.out_h, c.desc.out_w), -1.0, dtype=torch.float32, device="cuda:0") for c in cases]
    batch = capi.Batch(hip_ctx, [(c.desc, t[0], t[1], o) for c, t, o in zip(cases, ins, outs)], capture_graph=True)
    for _ in range(2):
        batch.run()
        batch.status()
        for c, a, b in zip(cases, want, outs):
            assert torch.equal(a, b), c.name
    batch.close()
