"""The cyclic-run query of the sorted sample order is a superset.

Both splat kernels (fg_pixelwise_hbm.hip, fg_grainwise.hip) walk, per grain,
only the samples whose offset fraction frac(X) lies in a widened cyclic
interval, looked up through a bucket index.  A sample outside that run must
never have an f32 candidate box with a pixel column.  This test evaluates the
box and the run in IEEE f32 (numpy, no fused operations) on random and
adversarial grain positions (centres at integers and half-integers, +-1e-4)
and checks that no sample with a non-empty box is left out.  It also checks
the run stays within a small factor of the true candidate count.
"""
import numpy as np
import pytest

f32 = np.float32


def frac(x):
    f = (x - np.floor(x)).astype(f32)
    return np.where(f < f32(1), f, f32(0)).astype(f32)


def run_positions(A, rz, eps, qidx, Q, N):
    """The kernels' run query for one grain: sorted positions walked."""
    lo = f32(f32(A - f32(rz + rz)) - eps)
    hi = f32(A + eps)
    Qf = f32(Q)
    if not (lo * Qf > -2 * Qf) or not (hi * Qf < 2 * Qf):
        return np.arange(N)
    lq, hq = int(np.floor(lo * Qf)), int(np.floor(hi * Qf)) + 1
    if hq - lq >= Q:
        return np.arange(N)
    if lq < 0:
        lq += Q
        hq += Q
    lq = min(max(lq, 0), Q - 1)
    hq = min(max(hq, lq + 1), 2 * Q)
    p0 = qidx[lq]
    end = qidx[hq] if hq <= Q else N + qidx[hq - Q]
    return (p0 + np.arange(end - p0)) % N


@pytest.mark.parametrize("N,z,sd", [(7, 3.0, 0.5), (64, 1.0, 0.8), (256, 2.0, 0.4), (1000, 1.7, 1.5)])
def test_run_contains_every_candidate(N, z, sd):
    rng = np.random.default_rng(N)
    o = rng.normal(0, sd, size=N).astype(f32)
    X = (o * f32(z)).astype(f32)
    f = frac(X)
    order = np.lexsort((np.arange(N), f))
    Xs = X[order]
    Q = 64
    while Q < N and Q < 4096:
        Q <<= 1
    qidx = np.searchsorted(f[order], (np.arange(Q + 1) / Q).astype(f32), side="left")
    out_w = 3840
    c_out = (4000 + 10) * float(z) + 8
    eps = f32(c_out * 2.0 ** -20 + 2.0 ** -16)
    M = 2000
    bx = (rng.uniform(-50, 4000, M).astype(f32) * f32(z) - f32(0.5)).astype(f32)
    bx[: M // 2] = (np.round(bx[: M // 2] * 2) / 2 + rng.choice([-1e-4, 0, 1e-4], M // 2)).astype(f32)
    rz = (rng.uniform(0.001, 0.8, M).astype(f32) * f32(z) + f32(4e-3 + 2e-6 * out_w)).astype(f32)
    walked = true = 0
    for b, r in zip(bx, rz):
        cxk = (b + Xs).astype(f32)
        has_col = np.ceil((cxk - r).astype(f32)) <= np.floor((cxk + r).astype(f32))
        nc = f32(r - b)
        A = f32(nc - np.floor(nc))
        sel = np.zeros(N, bool)
        sel[run_positions(A, r, eps, qidx, Q, N)] = True
        assert not np.any(has_col & ~sel), (N, z, float(b), float(r))
        walked += int(sel.sum())
        true += int(has_col.sum())
    assert walked <= 1.5 * true + 4 * M


def cell_members(fx, fy, qa, qb, Q, R, eps):
    """Membership of samples (fractions fx, fy) in the candidate list of
    cell (qa, qb), in the f32 operation order of cell_lists_kernel in
    fg_grainwise.hip."""
    m = f32(2.0 ** -12)
    w = f32(f32(f32(R + R) + f32(eps + eps)) + f32(f32(1) / f32(Q) + f32(m + m)))
    lx = f32(f32(f32(qa) / f32(Q) - f32(R + R)) - f32(eps + m))
    ly = f32(f32(f32(qb) / f32(Q) - f32(R + R)) - f32(eps + m))
    dx = (fx - lx).astype(f32)
    dy = (fy - ly).astype(f32)
    dx = (dx - np.floor(dx)).astype(f32)
    dy = (dy - np.floor(dy)).astype(f32)
    return (dx <= w) & (dy <= w)


@pytest.mark.parametrize("Q", [32, 64])
@pytest.mark.parametrize("N,R,sd", [(1, 0.05, 0.8), (128, 0.05, 0.8), (513, 0.159, 1.5), (2048, 0.01, 0.4)])
def test_cell_list_contains_every_candidate(N, R, sd, Q):
    """The grain-wise kernel's per-cell candidate lists (constant radius):
    every sample whose f32 bounds() box (src/grainwise.rs:126-142) has both
    a column and a row is in the list of the cell of the grain's target
    fractions (A, B), for grain centres anywhere in a 4096-pixel plane and
    at half-integers +- 1e-4."""
    rng = np.random.default_rng(N)
    X = rng.normal(0, sd, size=(N, 2)).astype(f32)
    fx = (X[:, 0] - np.floor(X[:, 0])).astype(f32)
    fy = (X[:, 1] - np.floor(X[:, 1])).astype(f32)
    R = f32(R)
    c = f32(f32(4096 + R + 8) + np.abs(X).max())
    eps = f32(f32(c * f32(2.0 ** -20)) + f32(2.0 ** -16))
    M = 3000
    b = rng.uniform(-8, 4100, size=(M, 2)).astype(f32)
    b[: M // 2] = (np.round(b[: M // 2] * 2) / 2 + rng.choice([-1e-4, 0, 1e-4], size=(M // 2, 2))).astype(f32)
    listed = true = 0
    for bx, by in b:
        tx = (bx + X[:, 0]).astype(f32)
        ty = (by + X[:, 1]).astype(f32)
        col = np.ceil(((tx - R).astype(f32) - f32(0.5)).astype(f32)) <= np.floor(
            ((tx + R).astype(f32) - f32(0.5)).astype(f32))
        row = np.ceil(((ty - R).astype(f32) - f32(0.5)).astype(f32)) <= np.floor(
            ((ty + R).astype(f32) - f32(0.5)).astype(f32))
        ncx, ncy = f32(R - f32(bx - f32(0.5))), f32(R - f32(by - f32(0.5)))
        A, B = f32(ncx - np.floor(ncx)), f32(ncy - np.floor(ncy))
        qa = min(max(int(f32(A * f32(Q))), 0), Q - 1)
        qb = min(max(int(f32(B * f32(Q))), 0), Q - 1)
        sel = cell_members(fx, fy, qa, qb, Q, R, eps)
        assert not np.any(col & row & ~sel), (N, float(R), float(bx), float(by))
        listed += int(sel.sum())
        true += int((col & row).sum())
    # about N (2R + 1/Q)^2 per cell against N (2R)^2 true candidates
    w = 2 * float(R) + 1.0 / Q + 2 * float(eps) + 2.0 ** -11
    assert listed <= 1.5 * M * N * w * w + 4 * M, (listed, true)


@pytest.mark.parametrize("z,out_side", [(0.5, 2048), (1.0, 640), (2.0, 3840), (3.7, 4096)])
def test_narrow_grain_passes_only_the_nearest_pixel(z, out_side):
    """The splat kernel's direct test (fg_pixelwise_hbm.hip): for a grain with
    2 rz < 1, rz = rad z + slack, a pixel whose point passes the reference's
    f32 disk test (src/pixelwise.rs) is the pixel nearest the sample's
    target cxk = (gx z - 0.5) + ox z, in both axes.  Checked in IEEE f32 on
    the 6x6 pixels around every target, for radii up to the largest narrow
    one and targets within 1e-3 of half-integers."""
    rng = np.random.default_rng(int(z * 10) + out_side)
    zf, iz = f32(z), f32(f32(1) / f32(z))
    slack = f32(f32(4e-3) + f32(f32(2e-6) * f32(out_side)))
    in_side = out_side / z
    M, N = 3000, 64
    g = rng.uniform(0, in_side, size=(M, 2)).astype(f32)
    o = rng.normal(0, 0.8, size=(N, 2)).astype(f32)
    r_max = float((f32(0.5) - slack) / zf)
    rad = rng.uniform(0.0, 1.0, M) ** 0.3 * r_max
    rad = np.minimum(rad, r_max * (1 - 1e-6)).astype(f32)
    # half the grains: targets of sample 0 within 1e-3 of a half-integer
    h = M // 2
    t = (np.floor(g[:h] * zf) + f32(0.5) + rng.uniform(-1e-3, 1e-3, size=(h, 2))).astype(f32)
    g[:h] = ((t - o[0] * zf + f32(0.5)) / zf).astype(f32)
    X = (o * zf).astype(f32)
    b = ((g * zf).astype(f32) - f32(0.5)).astype(f32)
    rz = (f32(1) * (rad * zf).astype(f32) + slack).astype(f32)
    assert np.all(rz + rz < f32(1))
    checked = hits = 0
    for i in range(M):
        ck = (b[i][None, :] + X).astype(f32)  # (N, 2) targets
        near = np.rint(ck).astype(f32)
        base = np.floor(ck).astype(np.int64)
        for dxp in range(-2, 4):
            for dyp in range(-2, 4):
                cx = (base[:, 0] + dxp).astype(f32)
                cy = (base[:, 1] + dyp).astype(f32)
                xg = (((cx + f32(0.5)).astype(f32) * iz).astype(f32) - o[:, 0]).astype(f32)
                yg = (((cy + f32(0.5)).astype(f32) * iz).astype(f32) - o[:, 1]).astype(f32)
                dx = (xg - g[i, 0]).astype(f32)
                dy = (yg - g[i, 1]).astype(f32)
                inside = ((dx * dx).astype(f32) + (dy * dy).astype(f32)).astype(f32) <= f32(rad[i] * rad[i])
                bad = inside & ((cx != near[:, 0]) | (cy != near[:, 1]))
                assert not np.any(bad), (z, out_side, i, float(rad[i]))
                hits += int(inside.sum())
                checked += N
    assert hits > 0


def splat_list_members(fx, fy, k, qa, qb, q, eps):
    """Membership of samples in the splat candidate list (class k, cell
    (qa, qb)), in the f32 operation order of splat_lists_kernel in
    fg_pixelwise_hbm.hip."""
    m = f32(2.0 ** -12)
    R = f32(f32(k + 1) / f32(2 * q))
    w = f32(f32(f32(R + R) + f32(eps + eps)) + f32(f32(f32(1) / f32(q)) + f32(m + m)))
    lx = f32(f32(f32(f32(qa) / f32(q)) - R) - f32(eps + m))
    ly = f32(f32(f32(f32(qb) / f32(q)) - R) - f32(eps + m))
    dx = (fx - lx).astype(f32)
    dy = (fy - ly).astype(f32)
    dx = (dx - np.floor(dx)).astype(f32)
    dy = (dy - np.floor(dy)).astype(f32)
    return (dx <= w) & (dy <= w)


def splat_list_span(k, q, e):
    return min(q, k + 4 + e)


@pytest.mark.parametrize("N,z,sd,out_side", [(100, 2.0, 0.4, 3840), (256, 2.0, 0.8, 3840), (513, 0.7, 1.5, 1024),
                                              (2048, 3.0, 0.8, 4096)])
def test_splat_list_contains_every_candidate(N, z, sd, out_side):
    """The splat's per-(radius class, cell) candidate lists: every sample
    whose f32 candidate box (cxk -+ rz, cxk = (cx z - 0.5) + ox z) has a
    column and a row is in the list of the grain's class floor(rz * 64) and
    of the cell of its centre fractions (frac(-bx), frac(-by)); and each
    class's lists fit the room splat_list_span gives them."""
    rng = np.random.default_rng(N + out_side)
    q = 32
    zf = f32(z)
    o = rng.normal(0, sd, size=(N, 2)).astype(f32)
    X = (o * zf).astype(f32)
    fx = (X[:, 0] - np.floor(X[:, 0])).astype(f32)
    fy = (X[:, 1] - np.floor(X[:, 1])).astype(f32)
    in_side = out_side / z
    c_in = in_side + 0.4 + float(np.abs(o).max()) + 4.0
    c_out = (c_in + float(np.abs(o).max()) + 0.4) * z + 8.0
    eps = f32(c_out * 2.0 ** -20 + 2.0 ** -16)
    e = int(np.ceil((2.0 * float(eps) + 2.0 ** -11) * q + 0.01))
    slack = f32(f32(4e-3) + f32(f32(2e-6) * f32(out_side)))
    M = 2500
    g = rng.uniform(-2, in_side + 2, size=(M, 2)).astype(f32)
    h = M // 2  # targets of sample 0 near half-integers
    t = (np.floor(g[:h] * zf) + f32(0.5) + rng.uniform(-1e-4, 1e-4, size=(h, 2))).astype(f32)
    g[:h] = ((t - X[0] + f32(0.5)) / zf).astype(f32)
    rad = (rng.uniform(0.0, 1.0, M) * (0.5 - float(slack)) / z).astype(f32)
    for i in range(M):
        bx = f32(f32(g[i, 0] * zf) - f32(0.5))
        by = f32(f32(g[i, 1] * zf) - f32(0.5))
        rz = f32(f32(rad[i] * zf) + slack)
        if not rz + rz < f32(1):
            continue
        cx = (bx + X[:, 0]).astype(f32)
        cy = (by + X[:, 1]).astype(f32)
        col = np.ceil((cx - rz).astype(f32)) <= np.floor((cx + rz).astype(f32))
        row = np.ceil((cy - rz).astype(f32)) <= np.floor((cy + rz).astype(f32))
        k = int(f32(rz * f32(2 * q)))
        ncx, ncy = f32(-bx), f32(-by)
        ax, ay = f32(ncx - np.floor(ncx)), f32(ncy - np.floor(ncy))
        qa, qb = min(int(f32(ax * f32(q))), q - 1), min(int(f32(ay * f32(q))), q - 1)
        sel = splat_list_members(fx, fy, k, qa, qb, q, eps)
        assert not np.any(col & row & ~sel), (N, z, i, float(rz))
    # the room per class: no list set exceeds n * span^2 entries
    for k in range(32):
        total = sum(int(splat_list_members(fx, fy, k, c % q, c // q, q, eps).sum()) for c in range(q * q))
        assert total <= N * splat_list_span(k, q, e) ** 2, (k, total)


def row_bucket(X_y, ext_z, ext_w, y):
    """Row bucket of y over the scaled y offsets (row_bucket in
    fg_pixelwise_hbm.hip), in f32."""
    inv_w = f32(f32(16) / f32(f32(ext_w - ext_z) + f32(0.015625)))
    b = np.floor((np.asarray(y, dtype=f32) - ext_z).astype(f32) * inv_w)
    return np.clip(b, 0, 15).astype(np.int64)


@pytest.mark.parametrize("N,z,sd,out_side,tile", [(256, 2.0, 0.8, 3840, 32), (100, 1.0, 0.4, 1920, 16),
                                                   (513, 0.7, 1.5, 1024, 64), (2048, 3.0, 0.8, 4096, 32)])
def test_splat_row_trim_keeps_every_tile_row(N, z, sd, out_side, tile):
    """The splat walks, of a grain's candidate list, only the row buckets of
    [Y0 - 0.5 - by, Y1 - 0.5 - by] widened by 2^-4, the list being ordered by
    X.y with per-bucket counts.  Every sample whose target row rint(fma(oy,
    z, by)) lies in the tile (the test's tile check, evaluated with the
    fused product) must have its bucket in that range; and the range drops
    the entries of rows far from the tile."""
    rng = np.random.default_rng(7 * N + tile)
    zf = f32(z)
    o = rng.normal(0, sd, size=(N, 2)).astype(f32)
    Xy = (o[:, 1] * zf).astype(f32)
    ext_z, ext_w = Xy.min(), Xy.max()
    bucket = row_bucket(Xy, ext_z, ext_w, Xy)
    order = np.lexsort((np.arange(N), Xy))
    # bucket counts of the (X.y, p) order are non-decreasing: the entries
    # of buckets b0..b1 are one run of the order
    assert np.all(np.diff(bucket[order]) >= 0)
    kept = total = 0
    tiles = out_side // tile
    for i in range(400):
        Y0 = int(rng.integers(0, tiles)) * tile
        Y1 = min(Y0 + tile, out_side)
        # grain centres over the rows whose samples can reach the tile, and
        # targets of sample 0 near the tile's edge rows
        cy_in = rng.uniform((Y0 - 12) / z, (Y1 + 12) / z)
        if i % 2:
            t = f32(Y0 - 0.5 + rng.choice([0.0, Y1 - Y0]) + rng.uniform(-1e-4, 1e-4))
            cy_in = float((t - Xy[0] + f32(0.5)) / zf)
        by = f32(f32(f32(cy_in) * zf) - f32(0.5))
        target = (o[:, 1].astype(np.float64) * float(zf) + float(by)).astype(f32)
        row = np.rint(target)
        in_tile = (row >= Y0) & (row <= Y1 - 1)
        fy0, fy1 = f32(Y0), f32(Y1 - 1)
        b0 = row_bucket(Xy, ext_z, ext_w, f32(f32(f32(fy0 - f32(0.5)) - by) - f32(0.0625)))
        b1 = row_bucket(Xy, ext_z, ext_w, f32(f32(f32(fy1 + f32(0.5)) - by) + f32(0.0625)))
        walked = (bucket >= b0) & (bucket <= b1)
        assert not np.any(in_tile & ~walked), (N, z, i, Y0, float(by))
        kept += int(walked.sum())
        total += N
    assert kept < total


def quarter_bucket_counts(member, jb, n):
    """The row-bucket counts of one cell as splat_lists_kernel writes them
    (fg_pixelwise_hbm.hip): four threads take the quarters of the (X.y,
    position) order; a bucket count is written by the quarter holding the
    bucket's first index, and buckets starting at n by quarter 0 with the
    list's length.  member: the cell's membership over the order; jb[b]: the
    first index of bucket >= b, b = 0..16."""
    qn = (n + 3) // 4
    counts = [member[q * qn:min(n, q * qn + qn)].sum() for q in range(4)]
    yx = [None] * 16
    length = int(sum(counts))
    for b in range(1, 17):
        if jb[b] == n:
            yx[b - 1] = length
    for q in range(4):
        p_lo, p_hi = q * qn, min(n, q * qn + qn)
        rel = int(sum(counts[:q]))
        b = 1
        while b <= 16 and jb[b] < p_lo:
            b += 1
        for i in range(p_lo, p_hi):
            while b <= 16 and jb[b] <= i:
                assert yx[b - 1] is None
                yx[b - 1] = rel
                b += 1
            rel += int(member[i])
    return yx


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 64, 129, 256, 2048])
def test_list_bucket_counts_equal_brute_force(n):
    """Every one of a cell's 16 counts is written exactly once and equals the
    number of its members in buckets below b, for bucket layouts with empty
    leading, inner and trailing buckets and quarters that are empty."""
    rng = np.random.default_rng(n)
    for trial in range(40):
        # non-decreasing buckets over the order, some buckets skipped
        bucket = np.sort(rng.integers(0, 16, size=n)) if trial % 4 else np.full(n, trial % 16)
        member = rng.random(n) < rng.choice([0.0, 0.3, 1.0])
        jb = [0] + [int(np.searchsorted(bucket, b, side="left")) for b in range(1, 16)] + [n]
        yx = quarter_bucket_counts(member, jb, n)
        want = [int(member[bucket < b].sum()) for b in range(1, 17)]
        assert yx == want, (n, trial)
