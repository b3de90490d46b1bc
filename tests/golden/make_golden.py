"""Writes tests/golden/fixtures.npz: small input/output vectors of the CPU
oracle (the restatement of the reference's CPU path in oracle/), at the
benchmark configurations' parameters.  The reference ships no fixtures or
tests of its own, so these are generated here, by the oracle only:

  c1_cells / c2_cells   per cell of a 16 x 16-cell patch: the cell_rng hash,
                        the seeded Xoshiro256++ state, the Poisson count and
                        up to 8 grains (cx, cy, r) (src/rng.rs:26-29,
                        src/pixelwise.rs:66-100), at C1 parameters (r = 0.12
                        const, delta = 1/9, seed 42) and C2 parameters
                        (lognormal 0.1 / 0.05, delta = 0.1, zoom 2, seed 7)
  c1_patch              a 32 x 32 output patch of the pixel-wise render at C1
                        parameters (N = 64)
  c3_bits               the grain-wise sample bitsets of a 16 x 16 output
                        patch at C3 parameters (r = 0.05 const, N = 128, seed 11)

Each carries the lambda plane and offsets it was computed from.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from tests import fgcases  # noqa: E402

OUT = Path(__file__).resolve().parent / "fixtures.npz"
MAX_G = 8


def cells(case, i0, j0, n=16):
    hash_ = np.zeros((n, n), dtype=np.uint64)
    state = np.zeros((n, n, 4), dtype=np.uint64)
    q = np.zeros((n, n), dtype=np.uint32)
    g = np.zeros((n, n, MAX_G, 3), dtype=np.float32)
    for dj in range(n):
        for di in range(n):
            h, s, qq, gg = O.cell_dump(case.plane, case.lam, i0 + di, j0 + dj, max_grains=MAX_G)
            hash_[dj, di], state[dj, di], q[dj, di] = h, s, qq
            g[dj, di, :len(gg)] = gg
    return hash_, state, q, g


def plane_fields(prefix, case):
    p = case.plane
    return {
        f"{prefix}_lam": case.lam, f"{prefix}_offsets_input": case.offsets_input, f"{prefix}_offsets": case.offsets,
        f"{prefix}_plane": np.array([p.in_w, p.in_h, p.out_w, p.out_h, p.n_samples, p.dist_kind], dtype=np.uint32),
        f"{prefix}_seed": np.array([p.seed], dtype=np.uint64),
        f"{prefix}_floats": np.array([p.zoom, p.delta, p.rm, p.radius_mean, p.log_mu, p.log_sigma], dtype=np.float32),
    }


def main():
    data = {}
    c1 = fgcases.c1_case(O, size=64, n_samples=64)
    data.update(plane_fields("c1", c1))
    h, s, q, g = cells(c1, 200, 240)
    data.update(c1_cells_origin=np.array([200, 240], dtype=np.int32), c1_cells_hash=h, c1_cells_state=s,
                c1_cells_q=q, c1_cells_grains=g)
    region = (16, 16, 48, 48)
    data.update(c1_patch_region=np.array(region, dtype=np.int32),
                c1_patch=O.render_pixelwise(c1.plane, c1.lam, c1.offsets_input, region=region))

    c2 = fgcases.c2_case(O, w=96, h=64, n_samples=256, channel=0)
    data.update(plane_fields("c2", c2))
    h, s, q, g = cells(c2, 300, 200)
    data.update(c2_cells_origin=np.array([300, 200], dtype=np.int32), c2_cells_hash=h, c2_cells_state=s,
                c2_cells_q=q, c2_cells_grains=g)

    c3 = fgcases.c3_case(O, size=48, n_samples=128)
    data.update(plane_fields("c3", c3))
    out, bits, pairs = O.render_grainwise(c3.plane, c3.lam, c3.offsets, want_bitsets=True)
    data.update(c3_bits_origin=np.array([16, 16], dtype=np.int32), c3_bits=bits[16:32, 16:32],
                c3_patch=out[16:32, 16:32], c3_pairs=np.array([pairs], dtype=np.uint64))
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes): {sorted(data)}")


if __name__ == "__main__":
    main()
