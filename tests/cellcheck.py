"""Cell-level comparison of the device generators with the oracle's.

Shared by the GPU cell tests and the golden-fixture tests: groups the grain
records of fg_pixelwise_grains by cell and compares counts and geometry with
fgo_cell_count / fgo_cell_dump; replays grain-wise pixel streams from the
device's post-count Xoshiro states with the oracle's own draw functions.
"""
from __future__ import annotations

import ctypes as C

import numpy as np


def group_cells(records: np.ndarray, origin, size):
    """(i, j) -> (k, 3) f32 grains in draw order, from (n, 4) records whose
    fourth float holds (i - i0) | (j - j0) << 16."""
    bits = records[:, 3].view(np.uint32)
    key = bits.astype(np.int64)
    change = np.flatnonzero(np.diff(key)) + 1
    starts = np.concatenate([[0], change]) if len(key) else np.zeros(0, dtype=np.int64)
    ends = np.concatenate([change, [len(key)]]) if len(key) else np.zeros(0, dtype=np.int64)
    cells = {}
    for a, b in zip(starts, ends):
        k = int(key[a])
        ij = (origin[0] + (k & 0xFFFF), origin[1] + (k >> 16))
        assert ij not in cells, f"cell {ij} split across the grain buffer"
        cells[ij] = records[a:b, :3]
    return cells


def ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """|a - b| in units in the last place of f32 (for finite values of one sign)."""
    ia = a.astype(np.float32).view(np.int32).astype(np.int64)
    ib = b.astype(np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def compare_cells(oracle, plane, lam, cells, origin, size, max_dump=4096):
    """Counts of every cell of the grid equal; returns (cells with grains,
    grains compared, grains not bit-equal, max ulp)."""
    n_cells, n_grains, inexact, worst = 0, 0, 0, 0
    W, H = size
    for i in range(origin[0], origin[0] + W):
        for j in range(origin[1], origin[1] + H):
            got = cells.get((i, j))
            q = oracle.cell_count(plane, lam, i, j)
            assert (0 if got is None else len(got)) == q, f"cell ({i}, {j}): device {0 if got is None else len(got)} grains, oracle {q}"
            if q == 0:
                continue
            _, _, q2, want = oracle.cell_dump(plane, lam, i, j, max_grains=min(q, max_dump))
            assert q2 == q
            d = ulp_diff(got[:len(want)], want)
            n_cells += 1
            n_grains += len(want)
            inexact += int(np.sum(d.max(axis=1) > 0))
            worst = max(worst, int(d.max()))
    return n_cells, n_grains, inexact, worst


def replay_pixel_grains(oracle, plane, x, y, state, q):
    """The grains of input pixel (x, y) that pixel_rng's stream yields after
    its count draws, drawn from `state` with the oracle's UniformFloat and
    LogNormal (src/grainwise.rs:46-58)."""
    L = oracle.lib()
    s = np.ascontiguousarray(np.array(state, dtype=np.uint64))
    sp = s.ctypes.data_as(C.POINTER(C.c_uint64))
    scale = L.fgo_uniform_f32_scale(0.0, 1.0)
    out = np.zeros((q, 3), dtype=np.float32)
    for g in range(q):
        out[g, 0] = np.float32(x) + np.float32(L.fgo_uniform_f32_sample(sp, 0.0, scale))
        out[g, 1] = np.float32(y) + np.float32(L.fgo_uniform_f32_sample(sp, 0.0, scale))
        if plane.dist_kind == 1:
            r = np.float32(L.fgo_lognormal_sample(sp, float(np.float32(plane.log_mu)), float(np.float32(plane.log_sigma))))
        else:
            r = np.float32(plane.radius_mean)
        out[g, 2] = min(r, np.float32(plane.rm))
    return out


def _splitmix_final(v):
    v = v.astype(np.uint64)
    with np.errstate(over="ignore"):
        v = v + np.uint64(0x9E3779B97F4A7C15)
        v = (v ^ (v >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        v = (v ^ (v >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return v ^ (v >> np.uint64(31))


def device_cell_digests(records: np.ndarray, size):
    """Per cell of the grid (h, w): the grain count and the digest the oracle's
    fgo_cell_digests computes, from (n, 4) records grouped by cell in draw order."""
    w, h = int(size[0]), int(size[1])
    q = np.zeros(w * h, dtype=np.uint32)
    d = np.zeros(w * h, dtype=np.uint64)
    n = len(records)
    if n == 0:
        return q.reshape(h, w), d.reshape(h, w)
    key = records[:, 3].view(np.uint32).astype(np.int64)
    starts = np.concatenate([[0], np.flatnonzero(np.diff(key)) + 1])
    lengths = np.diff(np.concatenate([starts, [n]]))
    k = np.arange(n, dtype=np.int64) - np.repeat(starts, lengths)
    b = records[:, :3].view(np.uint32).astype(np.uint64)
    with np.errstate(over="ignore"):
        v = (b[:, 0] << np.uint64(32)) | b[:, 1]
        v ^= b[:, 2] * np.uint64(0x9E3779B97F4A7C15)
        v ^= k.astype(np.uint64) * np.uint64(0xD1B54A32D192ED03)
    mixed = _splitmix_final(v)
    sums = np.add.reduceat(mixed, starts)
    cell = key[starts]
    idx = (cell >> 16) * w + (cell & 0xFFFF)
    assert len(np.unique(idx)) == len(idx), "a cell's grains are split across the buffer"
    q[idx] = lengths.astype(np.uint32)
    d[idx] = sums
    return q.reshape(h, w), d.reshape(h, w)
