"""CPU model of the deep fused sweep's index logic (overlapping windows of
64 lanes x 2 columns, D sweeps per pass, ring cells kept, guarded and steady
steps, an input ring of 3 to 12 rows, leading copy stages), checked against
D single sweeps of the oracle on chare blocks cut from a global grid.  The
model restates the device code's loop and window arithmetic step by step,
so a wrong window offset, ring column, store range, step split, input-ring
slot or band split shows up here before any GPU run."""
import numpy as np
import pytest

from oracle import oracle as O

K_COL_OFF = 15   # element offset of halo column 0 within a row
K_HALO = 16      # halo rows/columns every chare array holds
PLANT_NO_SELECTS = False  # a planted mistake: steady steps drop ring selects


def pitch_for_depth(w, d):
    return (w + 16 + 2 * d + 15) // 16 * 16


def shr1z(v):
    o = np.zeros_like(v)
    o[1:] = v[:-1]
    return o


def shl1z(v):
    o = np.zeros_like(v)
    o[:-1] = v[1:]
    return o


def window_margin(D):
    # D <= 8: 112 output columns (seven whole 128-byte lines); 9 <= D <= 16:
    # 96 (six)
    return 8 if D <= 8 else 16


def window_out_cols(D):
    return 128 - 2 * window_margin(D)


class Chare:
    """A chare array: rows 1-K_HALO .. bh+K_HALO, pitch_for_depth(bw, 16)."""

    def __init__(self, bw, bh):
        self.bw, self.bh = bw, bh
        self.pitch = pitch_for_depth(bw, K_HALO)
        self.a = np.zeros((bh + 2 * K_HALO, self.pitch))

    def idx(self, row, col):  # array index of (row, col), col may be array
        return row + K_HALO - 1, K_COL_OFF + col


def sweep_window(D, inp, out, w, h, c0, r0, r1, rs, acc, eo, pack, rin=3,
                 n_copy=0, ringc=0):
    """sweep_panelD<D, RES, PACK, 0, rin, ringc> on one band of one window,
    the first n_copy stages copying; ringc = 1 or 2: the window's ring
    columns fall only on the lanes' first or only on their second cells, and
    the steady steps keep them by selects."""
    assert rin in (3, 6, 9, 12)
    U = rin
    H = window_margin(D)
    lane = np.arange(64)
    c = c0 - H + 2 * lane
    vload = c <= w + D
    vstore = (lane >= H // 2) & (lane < 64 - H // 2)
    top, bottom, left, right = rs
    ring_a = (left & (c == 0)) | (right & (c == w + 1))
    ring_b = (left & (c + 1 == 0)) | (right & (c + 1 == w + 1))
    ringcols = bool(np.any(ring_a | ring_b))
    if ringc == 1:
        assert not np.any(ring_b)
    if ringc == 2:
        assert not np.any(ring_a)
    cend = w

    def load(r, row):
        # never past the band's input rows, never outside the array
        assert r0 - D <= row <= r1 + D - 1
        assert 1 - K_HALO <= row <= inp.bh + K_HALO
        i, j = inp.idx(row, c[vload])
        r[vload, 0] = inp.a[i, j]
        r[vload, 1] = inp.a[i, j + 1]

    def stage(up, cu, dn, q, check, copy=False, cols=0):
        if copy or (check and ((q == 0 and top) or (q == h + 1 and bottom))):
            return cu.copy()
        a, b = cu[:, 0], cu[:, 1]
        na = ((((a + shr1z(b)) + b) + up[:, 0]) + dn[:, 0]) * 0.2
        nb = ((((b + a) + shl1z(a)) + up[:, 1]) + dn[:, 1]) * 0.2
        if check and ringcols:
            na = np.where(ring_a, a, na)
            nb = np.where(ring_b, b, nb)
        elif not check and not PLANT_NO_SELECTS:
            if cols & 1:
                na = np.where(ring_a, a, na)
            if cols & 2:
                nb = np.where(ring_b, b, nb)
        return np.stack([na, nb], axis=1)

    def outrow(up, cu, dn, row):
        a, b = cu[:, 0], cu[:, 1]
        na = ((((a + shr1z(b)) + b) + up[:, 0]) + dn[:, 0]) * 0.2
        nb = ((((b + a) + shl1z(a)) + up[:, 1]) + dn[:, 1]) * 0.2
        for l in np.nonzero(vstore)[0]:
            cl = c[l]
            cells = [(cl, na[l], a[l])] if cl <= cend else []
            if cl + 1 <= cend:
                cells.append((cl + 1, nb[l], b[l]))
            for col, v, old in cells:
                i, j = out.idx(row, col)
                out.a[i, j] = v
                acc.append((v, old))
                if pack:
                    eo.record(row, col, v)

    R = [np.zeros((64, 2)) for _ in range(rin)]
    last_in = r1 + D - 1
    for k in range(rin):
        if r0 - D + k <= last_in:
            load(R[k], r0 - D + k)
    # NaN: a row read before it is made would reach the result
    S = [[np.full((64, 2), np.nan) for _ in range(3)] for _ in range(D - 1)]
    n_steps = r1 - r0 + 2 * D - 2
    tA = (2 * D - 2 + U - 1) // U * U
    ring_below = bottom and h + 1 <= r1 + D - 2
    t_hi = (min(n_steps - (D - 1), n_steps + 2 - rin) if ring_below
            else n_steps + 2 - rin)
    x = t_hi - tA
    tB = tA + int(x / U) * U   # C's division truncates toward zero
    if (not ringc and ringcols) or tB < tA:
        tB = tA
    t0 = 0
    while True:
        while t0 < n_steps and not (tA <= t0 < tB):
            for k in range(U):
                t, ks = t0 + k, k % 3
                if t < n_steps:
                    q1 = r0 - D + 1 + t
                    S[0][ks] = stage(R[k % rin], R[(k + 1) % rin],
                                     R[(k + 2) % rin], q1, True, n_copy >= 1)
                    if q1 - 1 + rin <= last_in:
                        load(R[k % rin], q1 - 1 + rin)
                    for s in range(2, D):
                        if t >= 2 * (s - 1):
                            S[s - 1][ks] = stage(S[s - 2][(ks + 1) % 3],
                                                 S[s - 2][(ks + 2) % 3],
                                                 S[s - 2][ks], q1 - (s - 1),
                                                 True, n_copy >= s)
                    if t >= 2 * (D - 1):
                        outrow(S[D - 2][(ks + 1) % 3], S[D - 2][(ks + 2) % 3],
                               S[D - 2][ks], q1 - (D - 1))
            t0 += U
        if t0 >= n_steps:
            break
        while t0 < tB:
            for k in range(U):
                ks = k % 3
                t = t0 + k
                q1 = r0 - D + 1 + t
                # the steady state must never meet what it skips
                assert not ((q1 == 0 and top) or (q1 == h + 1 and bottom))
                S[0][ks] = stage(R[k % rin], R[(k + 1) % rin], R[(k + 2) % rin],
                                 q1, False, n_copy >= 1, ringc)
                load(R[k % rin], q1 - 1 + rin)
                for s in range(2, D):
                    q = q1 - (s - 1)
                    assert not ((q == 0 and top) or (q == h + 1 and bottom))
                    S[s - 1][ks] = stage(S[s - 2][(ks + 1) % 3],
                                         S[s - 2][(ks + 2) % 3], S[s - 2][ks],
                                         q, False, n_copy >= s, ringc)
                outrow(S[D - 2][(ks + 1) % 3], S[D - 2][(ks + 2) % 3],
                       S[D - 2][ks], q1 - (D - 1))
            t0 += U


class Edges:
    """Outgoing edges D deep, as store_pair<D, RES, PACK> writes them."""

    def __init__(self, w, h, D):
        self.w, self.h, self.D = w, h, D
        self.top = np.full((D, w), np.nan)
        self.bottom = np.full((D, w), np.nan)
        self.left = np.full((D, h), np.nan)
        self.right = np.full((D, h), np.nan)

    def record(self, row, col, v):
        w, h, D = self.w, self.h, self.D
        if row <= D:
            self.top[row - 1, col - 1] = v
        if row >= h - D + 1:
            self.bottom[row - (h - D + 1), col - 1] = v
        if col <= D:
            self.left[col - 1, row - 1] = v
        if col >= w - D + 1:
            self.right[col - (w - D + 1), row - 1] = v


def ring_window(D, panel, bw, rs):
    """ring_window<D>: which cells of window `panel` are ring columns (0
    none, 1 only the lanes' first cells, 2 only their second cells, 3
    both)."""
    top, bottom, left, right = rs
    x0 = 1 + panel * window_out_cols(D) - window_margin(D)
    k = 0
    if left and x0 <= 0 <= x0 + 127:
        k |= 2 if (0 - x0) % 2 else 1
    if right and x0 <= bw + 1 <= x0 + 127:
        k |= 2 if (bw + 1 - x0) % 2 else 1
    return k


def chare_sweep(D, inp, out, rs, waves=8, band=0, rin=3, n_copy=0):
    """chare_sweepD: with band > 0, bands of `band` rows x windows taken
    round-robin, in three passes as band_tiles runs them (windows whose ring
    columns fall on the lanes' first cells, then on their second cells, each
    with a three-slot input ring and selects in the steady steps; then the
    others); else the windows x rows, window after window, cut into one
    contiguous range per wave (0: bands of 256 rows for chares of 1024+
    rows)."""
    bw, bh = inp.bw, inp.bh
    OW = window_out_cols(D)
    S = (bw + OW - 1) // OW
    if band == 0:
        band = 256 if bh >= 1024 else -1
    if band > 0:
        acc, eo = [], Edges(bw, bh, D)
        nb = (bh + band - 1) // band
        kinds = ([1, 2] if rs[2] or rs[3] else []) + [0]
        for kind in kinds:
            for t in range(S * nb):
                panel, b = t % S, t // S
                k = ring_window(D, panel, bw, rs)
                if (k in (1, 2)) if kind == 0 else (k != kind):
                    continue
                r0 = 1 + b * band
                sweep_window(D, inp, out, bw, bh, 1 + panel * OW, r0,
                             min(r0 + band, bh + 1), rs, acc, eo, True,
                             3 if kind else rin, n_copy, kind)
        return acc, eo
    total = S * bh
    per = (total + waves - 1) // waves
    acc, eo = [], Edges(bw, bh, D)
    for wave in range(waves):
        pos, end = wave * per, min(total, wave * per + per)
        while pos < end:
            panel, r = pos // bh, pos % bh
            n = min(bh - r, end - pos)
            sweep_window(D, inp, out, bw, bh, 1 + panel * OW, 1 + r,
                         1 + r + n, rs, acc, eo, True, rin, n_copy)
            pos += n
    return acc, eo


def cut(G, W, H, row0, col0, bw, bh):
    """A chare array filled from the global grid G the way the chares are
    seeded: every cell of the array with a global counterpart inside the
    ring gets it, the rest 0."""
    ch = Chare(bw, bh)
    for i in range(1 - K_HALO, bh + K_HALO + 1):
        gi = row0 + i
        if not 0 <= gi <= H + 1:
            continue
        lo, hi = max(1 - K_HALO, -col0), min(bw + K_HALO, W + 1 - col0)
        ii, jj = ch.idx(i, np.arange(lo, hi + 1))
        ch.a[ii, jj] = G[gi, col0 + lo: col0 + hi + 1]
    return ch


@pytest.mark.parametrize("rin", [3, 6, 9, 12])
@pytest.mark.parametrize("band", [0, 7, 64])
@pytest.mark.parametrize("D,n_copy", [(3, 0), (4, 0), (5, 0), (6, 0), (7, 0),
                                      (8, 0), (12, 0), (12, 3), (16, 0),
                                      (16, 1)])
@pytest.mark.parametrize("W,H,cr,cc,which", [
    (8, 8, 1, 1, (0, 0)), (12, 30, 2, 2, (1, 0)),
    (300, 200, 3, 5, (0, 0)), (300, 200, 3, 5, (1, 2)), (300, 200, 3, 5, (2, 4)),
    (260, 70, 1, 1, (0, 0)), (129, 257, 2, 2, (1, 1)), (1000, 40, 1, 3, (0, 2))])
def test_window_sweep_matches_oracle(D, n_copy, W, H, cr, cc, which, band,
                                     rin):
    G = O.init_grid(W, H, seed=W + 7 * H + D)
    ref = G.copy()
    for _ in range(D - n_copy):
        ref = O.step(ref)
    ci, cj = which
    # the last chare row/column takes the rest of an uneven split
    bh_base, bw_base = H // cr, W // cc
    row0 = ci * bh_base
    col0 = cj * bw_base
    bh = bh_base if ci < cr - 1 else H - bh_base * (cr - 1)
    bw = bw_base if cj < cc - 1 else W - bw_base * (cc - 1)
    if min(bw, bh) < D:
        pytest.skip("the runtime refuses chares thinner than the halo depth")
    inp = cut(G, W, H, row0, col0, bw, bh)
    out = Chare(bw, bh)
    rs = (ci == 0, ci == cr - 1, cj == 0, cj == cc - 1)
    acc, eo = chare_sweep(D, inp, out, rs, band=band, rin=rin, n_copy=n_copy)
    ii, jj = out.idx(np.arange(1, bh + 1)[:, None], np.arange(1, bw + 1)[None, :])
    got = out.a[ii, jj]
    want = ref[row0 + 1: row0 + bh + 1, col0 + 1: col0 + bw + 1]
    assert np.array_equal(got, want)
    # every interior cell stored exactly once
    assert len(acc) == bw * bh
    # the D-deep outgoing edges are the block's outer rows/columns
    assert np.array_equal(eo.top, want[:D, :])
    assert np.array_equal(eo.bottom, want[bh - D:, :])
    assert np.array_equal(eo.left, want[:, :D].T)
    assert np.array_equal(eo.right, want[:, bw - D:].T)


def test_model_catches_a_wrong_ring_column():
    # the left ring column sits in the second element of a lane's pair; a
    # model that only checks the first element must fail
    D, W, H = 3, 130, 40
    G = O.init_grid(W, H, seed=5)
    ref = G.copy()
    for _ in range(D):
        ref = O.step(ref)
    inp = cut(G, W, H, 0, 0, W, H)
    out = Chare(W, H)
    global sweep_window
    good = sweep_window

    def bad(D, inp, out, w, h, c0, r0, r1, rs, acc, eo, pack, *rest):
        top, bottom, left, right = rs
        return good(D, inp, out, w, h, c0, r0, r1, (top, bottom, False, right),
                    acc, eo, pack, *rest)
    sweep_window = bad
    try:
        chare_sweep(D, inp, out, (True, True, True, True))
    finally:
        sweep_window = good
    ii, jj = out.idx(np.arange(1, H + 1)[:, None], np.arange(1, W + 1)[None, :])
    assert not np.array_equal(out.a[ii, jj], ref[1:H + 1, 1:W + 1])




@pytest.mark.parametrize("D", [3, 8, 16])
def test_ring_windows_in_steady_steps(D):
    # bands with steady steps on a chare holding both ring columns (even and
    # odd widths put the right ring on either cell of a lane): the select
    # passes reproduce the oracle, and dropping the selects is caught
    global PLANT_NO_SELECTS
    for W in (300, 301):
        H = 90
        G = O.init_grid(W, H, seed=W + D)
        ref = G.copy()
        for _ in range(D):
            ref = O.step(ref)
        want = ref[1:H + 1, 1:W + 1]
        rs = (True, True, True, True)
        kinds = {ring_window(D, p, W, rs)
                 for p in range((W + window_out_cols(D) - 1) //
                                window_out_cols(D))}
        assert 2 in kinds and (1 in kinds) == (W % 2 == 0)
        for plant in (False, True):
            inp = cut(G, W, H, 0, 0, W, H)
            out = Chare(W, H)
            PLANT_NO_SELECTS = plant
            try:
                chare_sweep(D, inp, out, rs, band=40, rin=6)
            finally:
                PLANT_NO_SELECTS = False
            ii, jj = out.idx(np.arange(1, H + 1)[:, None],
                             np.arange(1, W + 1)[None, :])
            assert np.array_equal(out.a[ii, jj], want) != plant
