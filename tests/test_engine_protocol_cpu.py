"""CPU model of the task engine's dependency protocol (the batched and served
forms), run under random interleavings and checked against the oracle.

The model follows the device code of charming_amd/csrc/jacobi2d/batch.hip
step for step - per-tile counters for the next even/odd round, re-arming at
task start, up to eight neighbours, rounds of `depth` fused sweeps (eight
for a fixed sweep count, six with a tolerance, two for small tiles), the
stop-decision gate, the per-sweep residuals (exact in the first and last
round of a tolerance problem, an FP64 filter in between, with an exact
recompute of a round the filter cannot rule out while its last tile's
release waits), the repair of a stop inside a fused round, the in-flight
count and the state machine - while the stencil work is done with numpy on
the shared arrays.  Every task is split into a read phase and a write phase with other
tasks free to run in between, so two tasks the protocol lets overlap do
overlap.  If the protocol ever let a tile read cells a neighbour had not
finished, or overwrite cells a neighbour still needed, the result would
differ from the monolithic oracle's.
"""
import random
import sys

import numpy as np
import pytest

from oracle import oracle as O

TWO56 = float(2 ** 56)


def res_sum(new, old):
    """Exact sum of floor((new-old)^2 * 2^56), saturating like res_term."""
    d = new - old
    q = d * d * TWO56
    out = 0
    for v in q.ravel():
        out += (2 ** 64 - 1) if not v < 2.0 ** 64 else int(v)
    return out


def u128d(v):
    return O.u128_to_double(v)


def sweep_region(src, r0, r1, c0, c1, H, W):
    """One sweep of rows [r0, r1) x cols [c0, c1) of src (the stencil of
    jacobi2d.cu:332-333); cells on the grid's ring keep their value."""
    out = src[r0:r1, c0:c1].copy()
    i0, i1 = max(r0, 1), min(r1, H + 1)
    j0, j1 = max(c0, 1), min(c1, W + 1)
    if i0 < i1 and j0 < j1:
        c = src[i0:i1, j0:j1]
        v = (c + src[i0:i1, j0 - 1:j1 - 1] + src[i0:i1, j0 + 1:j1 + 1]
             + src[i0 - 1:i1 - 1, j0:j1] + src[i0 + 1:i1 + 1, j0:j1]) * 0.2
        out[i0 - r0:i1 - r0, j0 - c0:j1 - c0] = v
    return out


def depth_for(W, H, bs, tol2, deep=8):
    """Sweeps per round (build_request): `deep` (16 for engines of up to
    512-thread workgroups, else 8) for a fixed sweep count, 12 (engines of
    up to 512 threads) or 6 with a tolerance; 8 (6), then 2, when the tile
    or grid is thinner than that."""
    side = min(W, H, bs)
    if tol2 >= 0:
        deep_tol = 12 if deep == 16 else 6
        return deep_tol if side >= deep_tol else (6 if side >= 6 else 2)
    if side >= deep:
        return deep
    return 8 if side >= 8 else 2


class Problem:
    def __init__(self, W, H, bs, iters, tol, seed, deep=8):
        self.W, self.H, self.bs, self.iters = W, H, bs, iters
        self.tol2 = tol * tol if tol > 0 else -1.0
        self.tr, self.tc = -(-H // bs), -(-W // bs)
        self.n = self.tr * self.tc
        self.depth = depth_for(W, H, bs, self.tol2, deep)
        self.rounds = -(-iters // self.depth)
        g = O.init_grid(W, H, seed)
        self.A = [g.copy(), g.copy(), g.copy()]
        self.pend = []
        for u in range(self.n):
            deg = len(self.nbrs(u))
            self.pend.append([0, 1 + deg, 0])
        self.inflight = self.n
        self.state = 0
        self.fix_round = None
        self.fix_sweeps = 0
        self.round_tiles = [0, 0]
        self.res = [[0] * 16, [0] * 16]
        self.resd = [[0.0] * 16, [0.0] * 16]
        self.resid_round = None
        self.resid_tiles = 0
        self.held_tile = None
        self.s0 = 0
        self.stop_iters = 0
        self.s_last = 0
        self.hash = 0
        self.done = False

    def nbrs(self, u):
        ti, tj = divmod(u, self.tc)
        out = []
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                i, j = ti + di, tj + dj
                if (di or dj) and 0 <= i < self.tr and 0 <= j < self.tc:
                    out.append(i * self.tc + j)
        return out

    def geom(self, u):
        ti, tj = divmod(u, self.tc)
        r0, c0 = 1 + ti * self.bs, 1 + tj * self.bs
        return r0, min(r0 + self.bs, self.H + 1), c0, min(c0 + self.bs, self.W + 1)


def filtered_round(p, k):
    return p.tol2 >= 0 and p.depth >= 3 and 0 < k < p.rounds - 1


def filter_may_stop(p, sk, n_sw):
    """The device's bound: A 2^56 (1 - 2^-20) - N > thr (1 + 2^-20) proves
    that the exact residual exceeds thr."""
    thr = p.tol2 * u128d(p.s0)
    if not thr < 2.0 ** 62:
        return True
    n = float(p.W * p.H)
    for j in range(n_sw):
        a = p.resd[sk][j]
        if not (a * 2.0 ** 56 * (1 - 2.0 ** -20) - n > thr * (1 + 2.0 ** -20)):
            return True
    return False


# force the exact recompute of every filtered round (set by a test)
ALWAYS_RECOMPUTE = False


class Engine:
    def __init__(self, probs, rng):
        self.probs = probs
        self.rng = rng
        self.queue = []       # ready task words
        self.running = []     # (task, pending write)
        for pi, p in enumerate(probs):
            for u in range(p.n):
                self.queue.append(("iter", pi, u))

    def push(self, t):
        self.queue.append(t)

    # ---- task start: the counter re-arm ----
    def start(self, t):
        kind, pi, u = t
        p = self.probs[pi]
        if kind == "iter":
            k = p.pend[u][2]
            gate = 1 if p.tol2 >= 0 else 0  # round k+2 waits for decision k
            p.pend[u][k & 1] = 1 + len(p.nbrs(u)) + gate
            return k
        if kind == "resid":
            return p.resid_round
        return p.fix_round if kind == "fix" else None

    @staticmethod
    def fused(p, src, n_sw, r0, r1, c0, c1):
        """n_sw sweeps of the tile in one pass: sweep s over the tile grown
        by n_sw - s cells (what the pass reads), every sweep's residual over
        the tile's own cells."""
        big = src.copy()
        res = []
        for s in range(1, n_sw + 1):
            e = n_sw - s
            a0, a1 = max(r0 - e, 0), min(r1 + e, p.H + 2)
            b0, b1 = max(c0 - e, 0), min(c1 + e, p.W + 2)
            new = sweep_region(big, a0, a1, b0, b1, p.H, p.W)
            tile = new[r0 - a0:r1 - a0, c0 - b0:c1 - b0]
            res.append(res_sum(tile, big[r0:r1, c0:c1]))
            big[a0:a1, b0:b1] = new
        return big[r0:r1, c0:c1].copy(), res

    @staticmethod
    def fused_sq(p, src, n_sw, r0, r1, c0, c1):
        """Like fused, the FP64 sums of squared changes of every sweep."""
        big = src.copy()
        out = []
        for s in range(1, n_sw + 1):
            e = n_sw - s
            a0, a1 = max(r0 - e, 0), min(r1 + e, p.H + 2)
            b0, b1 = max(c0 - e, 0), min(c1 + e, p.W + 2)
            new = sweep_region(big, a0, a1, b0, b1, p.H, p.W)
            d = new[r0 - a0:r1 - a0, c0 - b0:c1 - b0] - big[r0:r1, c0:c1]
            out.append(float(np.sum(d * d)))
            big[a0:a1, b0:b1] = new
        return out

    # ---- the read phase: at the task's start or just before its end ----
    def read(self, t, k):
        kind, pi, u = t
        p = self.probs[pi]
        r0, r1, c0, c1 = p.geom(u)
        if kind == "iter":
            n_sw = min(p.depth, p.iters - p.depth * k)
            tol, last = p.tol2 >= 0, k + 1 == p.rounds
            new, res = self.fused(p, p.A[k % 3], n_sw, r0, r1, c0, c1)
            # what the device adds: every sweep's residual with a tolerance
            # (FP64 sums of squared changes in filtered rounds); on the last
            # round of a fixed count the last one, into the second slot for
            # a pass of two or more sweeps
            if tol and filtered_round(p, k):
                add = ("approx", self.fused_sq(p, p.A[k % 3], n_sw, r0, r1,
                                               c0, c1))
            elif tol:
                add = res
            elif last:
                add = [res[-1]] if n_sw == 1 else [0, res[-1]]
            else:
                add = []
            return (k, new, add)
        if kind == "fix":
            new, _ = self.fused(p, p.A[k % 3], p.fix_sweeps, r0, r1, c0, c1)
            return (k, new, [])
        if kind == "resid":
            n_sw = min(p.depth, p.iters - p.depth * k)
            new, res = self.fused(p, p.A[k % 3], n_sw, r0, r1, c0, c1)
            return (k, new, res)
        res_arr = ((p.stop_iters + p.depth - 1) // p.depth) % 3
        return (None, O.grid_hash(p.A[res_arr], r0, r1, c0, c1), [])

    # ---- task end: the write phase and the completion logic ----
    def finish(self, t, data):
        kind, pi, u = t
        p = self.probs[pi]
        r0, r1, c0, c1 = p.geom(u)
        k, new, add = data
        if kind == "iter":
            p.A[(k + 1) % 3][r0:r1, c0:c1] = new
            p.pend[u][2] = k + 1
            flags = self.complete_round(p, pi, u, k, add)
            if flags == "gate":
                self.gate(p, pi, k)
        elif kind == "fix":
            p.A[(k + 1) % 3][r0:r1, c0:c1] = new
        elif kind == "resid":
            # the recompute writes what the round wrote
            assert np.array_equal(p.A[(k + 1) % 3][r0:r1, c0:c1], new)
            p.A[(k + 1) % 3][r0:r1, c0:c1] = new
            sk = k & 1
            for j, v in enumerate(add):
                p.res[sk][j] += v
            p.resid_tiles += 1
            if p.resid_tiles == p.n:
                flags = self.decide_exact(p, k, sk)
                if p.state == 0:
                    self.release(p, pi, p.held_tile, k)
                    if flags == "gate":
                        self.gate(p, pi, k)
        else:
            p.hash = (p.hash + new) % 2 ** 64
        self.finish_task(p, pi)

    def gate(self, p, pi, k):
        """The decision on round k lets round k+2 of every tile go."""
        if k + 2 < p.rounds:
            for v in range(p.n):
                p.pend[v][k & 1] -= 1
                if p.pend[v][k & 1] == 0:
                    p.inflight += 1
                    self.push(("iter", pi, v))

    def release(self, p, pi, u, k):
        """Tile u's round k is done: its and its neighbours' round k+1."""
        if k + 1 < p.rounds:
            for v in [u] + p.nbrs(u):
                p.pend[v][(k + 1) & 1] -= 1
                assert p.pend[v][(k + 1) & 1] >= 0
                if p.pend[v][(k + 1) & 1] == 0:
                    p.inflight += 1
                    self.push(("iter", pi, v))

    def complete_round(self, p, pi, u, k, add):
        flags = self.decide(p, k, add)
        if flags == "resid":
            # every tile recomputes round k exactly; u's release waits
            p.held_tile = u
            for v in range(p.n):
                self.push(("resid", pi, v))
            return None
        self.release(p, pi, u, k)
        return flags

    def decide(self, p, k, add):
        tol, last = p.tol2 >= 0, k + 1 == p.rounds
        if not tol and not last:
            return None
        sk = (k & 1) if tol else 0
        if isinstance(add, tuple):
            for j, v in enumerate(add[1]):
                p.resd[sk][j] += v
        else:
            for j, v in enumerate(add):
                p.res[sk][j] += v
        p.round_tiles[sk] += 1
        if p.round_tiles[sk] != p.n:
            return None
        if p.state != 0:
            return None  # a stop was decided on the previous round
        if filtered_round(p, k):
            if ALWAYS_RECOMPUTE or filter_may_stop(p, sk, p.depth):
                p.resid_round, p.resid_tiles, p.state = k, 0, 5
                p.inflight += p.n
                return "resid"
            p.round_tiles[sk] = 0
            p.resd[sk] = [0.0] * 16
            return "gate"
        return self.decide_exact(p, k, sk)

    def decide_exact(self, p, k, sk):
        tol, last = p.tol2 >= 0, k + 1 == p.rounds
        n_sw = min(p.depth, p.iters - p.depth * k)
        stop = 0
        if not tol:
            stop, s = p.iters, p.res[sk][1 if n_sw >= 2 else 0]
        else:
            if k == 0:
                p.s0 = p.res[sk][0]
            thr = p.tol2 * u128d(p.s0)
            for j in range(n_sw):
                if u128d(p.res[sk][j]) <= thr or (last and j == n_sw - 1):
                    stop, s = p.depth * k + j + 1, p.res[sk][j]
                    if j + 1 < n_sw:
                        p.fix_round, p.fix_sweeps = k, j + 1
                    break
        if stop:
            p.s_last, p.stop_iters, p.state = s, stop, 1
            return None
        p.round_tiles[sk] = 0
        p.res[sk] = [0] * 16
        p.resd[sk] = [0.0] * 16
        p.state = 0
        return "gate"

    def finish_task(self, p, pi):
        p.inflight -= 1
        assert p.inflight >= 0
        if p.inflight:
            return
        if p.state in (1, 4):
            fix = p.state == 1 and p.fix_round is not None
            p.state = 4 if fix else 2
            p.inflight += p.n
            for v in range(p.n):
                self.push(("fix" if fix else "hash", pi, v))
        elif p.state == 2:
            p.done = True
        else:
            raise AssertionError("problem drained while running")

    def run(self, max_overlap):
        while self.queue or self.running:
            start = self.queue and (not self.running or
                                    (len(self.running) < max_overlap and
                                     self.rng.random() < 0.6))
            if start:
                t = self.queue.pop(self.rng.randrange(len(self.queue)))
                k = self.start(t)
                early = self.rng.random() < 0.5
                self.running.append((t, k, self.read(t, k) if early else None))
            else:
                t, k, data = self.running.pop(self.rng.randrange(len(self.running)))
                self.finish(t, data if data is not None else self.read(t, k))
        assert all(p.done for p in self.probs)


CASES = [  # W, H, tile, sweeps, tol
    (40, 30, 16, 9, 0.0), (33, 17, 16, 6, 0.0), (20, 20, 8, 1, 0.0),
    (50, 40, 16, 400, 3e-2), (24, 24, 8, 400, 0.5), (30, 18, 8, 7, 1e-9),
    (17, 33, 16, 400, 1e-2), (12, 12, 4, 2, 0.0), (40, 40, 16, 29, 0.0),
    (26, 19, 8, 400, 2e-2), (45, 21, 8, 400, 7e-3), (20, 20, 4, 15, 0.0),
    (24, 24, 8, 400, 5e-2), (30, 22, 8, 400, 0.3),
    (12, 12, 4, 160, 1e-13)]   # converges below the filter's bound


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_protocol_matches_oracle_under_random_interleavings(seed):
    rng = random.Random(seed)
    probs = [Problem(W, H, bs, it, tol, 100 * seed + i)
             for i, (W, H, bs, it, tol) in enumerate(CASES)]
    Engine(probs, rng).run(max_overlap=12)
    for (W, H, bs, it, tol), p in zip(CASES, probs):
        s = 100 * seed + CASES.index((W, H, bs, it, tol))
        if tol > 0:
            g, n, s_last, s0 = O.jacobi_solve_tol(W, H, it, tol, seed=s)
        else:
            (g, s_last), n = O.jacobi_run(W, H, it, seed=s), it
        assert p.stop_iters == n, (W, H, bs, it, tol)
        assert p.hash == O.grid_hash(g), (W, H, bs, it, tol)
        assert p.s_last == s_last, (W, H, bs, it, tol)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_protocol_with_sixteen_deep_rounds(seed):
    # fixed sweep counts in rounds of sixteen (tiles and grids of 16 cells
    # or more), eight or two below that, under random interleavings
    cases = [(40, 30, 16, 9, 0.0), (33, 17, 16, 37, 0.0), (40, 40, 16, 29, 0.0),
             (64, 48, 16, 50, 0.0), (20, 20, 8, 17, 0.0), (12, 12, 4, 5, 0.0),
             (40, 40, 20, 16, 0.0), (50, 40, 16, 400, 3e-2),
             (24, 24, 12, 400, 0.5), (26, 19, 13, 400, 2e-2),
             (45, 21, 12, 400, 7e-3), (30, 22, 12, 400, 0.3),
             (24, 24, 8, 400, 5e-2), (12, 12, 12, 160, 1e-13)]
    rng = random.Random(seed)
    probs = [Problem(W, H, bs, it, tol, 700 * seed + i, deep=16)
             for i, (W, H, bs, it, tol) in enumerate(cases)]
    assert {p.depth for p in probs} == {16, 12, 8, 6, 2}
    Engine(probs, rng).run(max_overlap=12)
    for i, ((W, H, bs, it, tol), p) in enumerate(zip(cases, probs)):
        s = 700 * seed + i
        if tol > 0:
            g, n, s_last, s0 = O.jacobi_solve_tol(W, H, it, tol, seed=s)
        else:
            (g, s_last), n = O.jacobi_run(W, H, it, seed=s), it
        assert (p.stop_iters, p.hash, p.s_last) == (n, O.grid_hash(g), s_last)


@pytest.mark.parametrize("deep", [8, 16])
def test_protocol_with_every_filtered_round_recomputed(monkeypatch, deep):
    # the exact recompute, the held release and the deferred gate on every
    # filtered round (not only where a stop may lie) still give the oracle's
    # answers, in rounds of six and of twelve
    monkeypatch.setattr(sys.modules[__name__], "ALWAYS_RECOMPUTE", True)
    rng = random.Random(11)
    cases = [c for c in CASES if c[4] > 0]
    probs = [Problem(W, H, bs, it, tol, 500 + i, deep=deep)
             for i, (W, H, bs, it, tol) in enumerate(cases)]
    Engine(probs, rng).run(max_overlap=12)
    for i, ((W, H, bs, it, tol), p) in enumerate(zip(cases, probs)):
        g, n, s_last, s0 = O.jacobi_solve_tol(W, H, it, tol, seed=500 + i)
        assert (p.stop_iters, p.hash, p.s_last) == (n, O.grid_hash(g), s_last)


def test_model_detects_a_missing_corner_dependency(monkeypatch):
    # the check has teeth: without the diagonal neighbours a fused round can
    # read a corner cell before (or after) its tile wrote it
    def sides_only(self, u):
        ti, tj = divmod(u, self.tc)
        out = []
        for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            i, j = ti + di, tj + dj
            if 0 <= i < self.tr and 0 <= j < self.tc:
                out.append(i * self.tc + j)
        return out
    monkeypatch.setattr(Problem, "nbrs", sides_only)
    bad = 0
    for seed in range(6):
        rng = random.Random(seed)
        p = Problem(40, 40, 8, 12, 0.0, seed)
        Engine([p], rng).run(max_overlap=16)
        g, _ = O.jacobi_run(40, 40, 12, seed=seed)
        bad += p.hash != O.grid_hash(g)
    assert bad > 0
