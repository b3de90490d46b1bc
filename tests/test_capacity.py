"""The grain-capacity limit of the pixel-wise HBM path, on the CPU.

The count pass sums each bin's grains in 64 bits (bin_count_planes_kernel) and the
host refuses a plane whose total exceeds fg_max_grains() before bin_grains
writes anything, because bin offsets, grain indices and the splat's chunk
indices are u32.  These tests check that limit and show, with the oracle's own
per-cell Poisson counts (src/pixelwise.rs:69-82), that reachable parameters
pass 2^32 grains while the wrapped u32 sum would look small.  The overflowing
case is never rendered on a GPU.
"""
import numpy as np

from film_grain_amd import capi


def test_limit_leaves_room_for_chunk_indices():
    lim = int(capi.lib().fg_max_grains())
    # the splat reads chunk indices up to 63 past the last grain; they must
    # not wrap a u32
    assert 0 < lim and lim + 64 < 2 ** 32


def test_large_cells_pass_2_pow_32_grains_and_u32_sum_wraps(oracle):
    # A bright 8x8 plane (lambda = 1e6, the MAX_LAMBDA cap) with a 64-px cell:
    # every cell's expected count is 1e6 * 64 * 64 = 4.1e9, so two cells
    # already hold more than 2^32 grains.
    O = oracle
    lam = np.full((8, 8), 1.0e6, dtype=np.float32)
    d = O.derive(0.05, 0.0, "const", 1.0, 0.8, 4, cell_delta=64.0, seed=3, in_w=8, in_h=8)
    p = O.plane(d, 4, 3, 1.0, "const", 0.05)
    counts = []
    for i in (0, -1):
        for j in (0, -1):
            counts.append(O.cell_count(p, lam, i, j))
    assert all(q > 4.0e9 for q in counts)
    total = sum(counts)
    lim = int(capi.lib().fg_max_grains())
    assert total > lim
    # a u32 running sum would wrap to a value the limit check lets through
    assert total % (2 ** 32) < lim
