"""The cell-comparison helpers the GPU cell tests use, checked on the CPU:
records built from the oracle's own dumps pass, and planted slips (a dropped
grain, a swapped pair, a one-ulp radius, a wrong count) are caught."""
import numpy as np
import pytest

from tests import cellcheck, fgcases


def _records(O, case, i0, j0, W, H):
    rows = []
    for j in range(j0, j0 + H):  # any cell order: grouping keys on the cell field
        for i in range(i0, i0 + W):
            _, _, q, g = O.cell_dump(case.plane, case.lam, i, j, max_grains=4096)
            bits = np.uint32((i - i0) | ((j - j0) << 16))
            for cx, cy, r in g:
                rows.append([cx, cy, r, bits.view(np.float32)])
    return np.array(rows, dtype=np.float32).reshape(-1, 4)


@pytest.fixture(scope="module")
def setup(oracle):
    case = fgcases.c2_case(oracle, w=6, h=4, n_samples=8)
    i0, j0, W, H = -3, -2, 66, 46
    return case, _records(oracle, case, i0, j0, W, H), (i0, j0), (W, H)


def test_oracle_records_pass(oracle, setup):
    case, rec, origin, size = setup
    cells = cellcheck.group_cells(rec, origin, size)
    n_cells, n_grains, inexact, worst = cellcheck.compare_cells(oracle, case.plane, case.lam, cells, origin, size)
    assert n_grains == len(rec) > 50 and inexact == 0 and worst == 0


def test_planted_slips_fail(oracle, setup):
    case, rec, origin, size = setup
    dropped = np.delete(rec, 7, axis=0)
    with pytest.raises(AssertionError):
        cellcheck.compare_cells(oracle, case.plane, case.lam, cellcheck.group_cells(dropped, origin, size),
                                origin, size)
    # two grains of one cell swapped (same cell field, different draw order)
    key = rec[:, 3].view(np.uint32)
    k = int(np.flatnonzero(key[1:] == key[:-1])[0])
    swapped = rec.copy()
    swapped[[k, k + 1]] = swapped[[k + 1, k]]
    cells = cellcheck.group_cells(swapped, origin, size)
    assert cellcheck.compare_cells(oracle, case.plane, case.lam, cells, origin, size)[2] > 0
    nudged = rec.copy()
    nudged[3, 2] = np.nextafter(nudged[3, 2], np.float32(1))
    cells = cellcheck.group_cells(nudged, origin, size)
    _, _, inexact, worst = cellcheck.compare_cells(oracle, case.plane, case.lam, cells, origin, size)
    assert inexact == 1 and worst == 1


def test_device_digests_equal_oracle_digests(oracle, setup):
    case, rec, origin, size = setup
    q, d = cellcheck.device_cell_digests(rec, size)
    qo, do = oracle.cell_digests(case.plane, case.lam, origin, size)
    assert np.array_equal(q, qo) and np.array_equal(d, do)
    key = rec[:, 3].view(np.uint32)
    k = int(np.flatnonzero(key[1:] == key[:-1])[0])
    swapped = rec.copy()
    swapped[[k, k + 1]] = swapped[[k + 1, k]]
    assert not np.array_equal(cellcheck.device_cell_digests(swapped, size)[1], do)
    nudged = rec.copy()
    nudged[3, 2] = np.nextafter(nudged[3, 2], np.float32(1))
    assert not np.array_equal(cellcheck.device_cell_digests(nudged, size)[1], do)


def test_oracle_pixel_counts_equal_pixel_dumps(oracle):
    case = fgcases.c3_case(oracle, size=16, n_samples=8)
    q, st = oracle.pixel_counts(case.plane, case.lam, states=True)
    for y in range(16):
        for x in range(16):
            _, _, qd, g = oracle.pixel_dump(case.plane, case.lam, x, y)
            assert q[y, x] == qd
            if qd:
                assert np.array_equal(cellcheck.replay_pixel_grains(oracle, case.plane, x, y, st[y, x], qd), g)
