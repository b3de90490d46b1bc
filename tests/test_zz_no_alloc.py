"""The allocation-free pixel-wise entry point, on MI355X, through the C ABI.

fg_render_pixelwise_device_noalloc takes every scratch buffer from the
library's static device array (256 MiB per device) instead of allocating.
Its planes must equal fg_render_pixelwise_device's bit for bit, on two calls
in a row (the second with a smaller plane, so it reuses the array that still
holds the first call's contents), device memory must not change across a
call, and a plane beyond the array is refused with FG_ECAPACITY, after which
the entry point still works.
"""
import numpy as np
import pytest

from film_grain_amd import capi
from tests import fgcases

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.device("cuda:0"))


def _pair(ref_ctx, na_ctx, case):
    import torch
    lam, offs = _dev(case.lam), _dev(case.offsets_input)
    want = torch.empty((case.desc.out_h, case.desc.out_w), dtype=torch.float32, device=lam.device)
    got = torch.full_like(want, -1.0)
    ref_ctx.render_device("pixel", case.desc, lam, offs, want)
    ref_ctx.stream_status()
    free0 = torch.cuda.mem_get_info()[0]
    na_ctx.render_device_noalloc(case.desc, lam, offs, got)
    na_ctx.stream_status()
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    return want.cpu().numpy(), got.cpu().numpy(), free0, free1


def test_noalloc_equals_device_entry_point(hip_ctx, oracle):
    import torch
    na = capi.Context(0)
    try:
        big = fgcases.c1_case(oracle, size=256, n_samples=64)
        small = fgcases.c2_case(oracle, w=48, h=32, n_samples=256, channel=1)
        print(f"[noalloc] {big.name} then {small.name}", flush=True)
        want, got, _, _ = _pair(hip_ctx, na, big)
        assert np.array_equal(got, want), f"{big.name}: max |d| = {np.abs(got - want).max()}"
        want, got, f0, f1 = _pair(hip_ctx, na, small)
        assert np.array_equal(got, want), f"{small.name}: max |d| = {np.abs(got - want).max()}"
        assert f1 == f0, f"device memory changed across the call: {f0} -> {f1}"
        # the small plane also equals the oracle
        assert np.array_equal(got, oracle.render_pixelwise(small.plane, small.lam, small.offsets_input))
        # a plane beyond the array: refused, then the entry point still works
        huge = fgcases.c2_case(oracle, w=1920, h=1080, n_samples=256, channel=0)
        print(f"[noalloc] {huge.name}: expecting FG_ECAPACITY", flush=True)
        out = torch.empty((huge.desc.out_h, huge.desc.out_w), dtype=torch.float32, device=torch.device("cuda:0"))
        with pytest.raises(capi.FilmGrainError) as err:
            na.render_device_noalloc(huge.desc, _dev(huge.lam), _dev(huge.offsets_input), out)
        assert err.value.code == capi.FG_ECAPACITY and "MiB" in str(err.value)
        na.stream_status()
        want, got, _, _ = _pair(hip_ctx, na, small)
        assert np.array_equal(got, want)
    finally:
        na.close()
