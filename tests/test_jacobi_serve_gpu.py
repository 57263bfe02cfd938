"""Served form: requests stream into the running engine (continuous batching,
admission against an HBM budget) or into the plain per-request form.  Each
result must equal the CPU oracle's and the same request solved alone."""
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cam():
    import charming_amd
    charming_amd.jacobi_lib()
    return charming_amd


def drain(srv, tickets, timeout=60.0):
    got = {}
    t0 = time.time()
    while len(got) < len(tickets):
        for c in srv.poll():
            got[c["ticket"]] = c
        assert time.time() - t0 < timeout, f"{len(got)}/{len(tickets)} done"
        time.sleep(0.0005)
    return got


def requests(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append({"width": int(rng.integers(8, 260)),
                    "height": int(rng.integers(8, 260)),
                    "iters": int(rng.integers(5, 300)),
                    "tol": float(rng.choice([1e-2, 3e-2, 1e-3, 1e-9])),
                    "seed": seed * 100 + i,
                    "block": int(rng.choice([0, 32, 64]))})
    return out


def oracle_tol(p):
    g, n, s_last, s0 = O.jacobi_solve_tol(p["width"], p["height"], p["iters"],
                                          p["tol"], seed=p["seed"])
    return O.grid_hash(g), n, s_last, s0


@pytest.mark.parametrize("mode", ["engine", "plain"])
def test_stream_matches_oracle(cam, mode):
    reqs = requests(24, 1)
    print("server", mode, flush=True)
    with cam.Server(mode, budget_bytes=1 << 30) as srv:
        tickets = {}
        for p in reqs:
            t = srv.submit(p)
            assert t is not None
            tickets[t] = p
        got = drain(srv, tickets)
    for t, p in tickets.items():
        h, n, s_last, s0 = oracle_tol(p)
        c = got[t]
        assert (c["hash"], c["iters"], c["residual"], c["residual0"]) == \
            (h, n, s_last, s0), p
        assert c["latency_ms"] >= 0 and c["device_ms"] >= 0


@pytest.mark.parametrize("threads", [512, 256])
def test_twelve_deep_tolerance_rounds_match_oracle(cam, threads):
    # engines of up to 512-thread workgroups run tolerance requests in rounds
    # of twelve fused sweeps (six, then two, on thinner grids and tiles): stops
    # inside a round are repaired, filtered rounds recomputed where a stop
    # may lie; fixed-count requests go in rounds of sixteen
    reqs = requests(20, 7 + threads)
    for i, p in enumerate(reqs):
        p["block"] = [0, 16, 24, 96, 192][i % 5]
        p["iters"] = 20 + 37 * i
    reqs.append({"width": 300, "height": 280, "iters": 77, "seed": 901})
    print("server", threads, flush=True)
    with cam.Server("engine", budget_bytes=1 << 30, wg_threads=threads) as srv:
        tickets = {}
        for p in reqs:
            t = srv.submit(p)
            assert t is not None
            tickets[t] = p
        got = drain(srv, tickets)
    for t, p in tickets.items():
        c = got[t]
        if "tol" in p:
            h, n, s_last, s0 = oracle_tol(p)
            assert (c["hash"], c["iters"], c["residual"], c["residual0"]) == \
                (h, n, s_last, s0), p
        else:
            g, res = O.jacobi_run(p["width"], p["height"], p["iters"],
                                  seed=p["seed"])
            assert (c["hash"], c["iters"], c["residual"]) == \
                (O.grid_hash(g), p["iters"], res), p


def test_result_alone_equals_in_mix(cam):
    reqs = requests(16, 2)
    with cam.Server("engine", budget_bytes=1 << 30) as srv:
        mix = drain(srv, {srv.submit(p): i for i, p in enumerate(reqs)})
        mix = {c["ticket"]: c for c in mix.values()}
        for i in (0, 5, 11):
            t = srv.submit(reqs[i])
            alone = drain(srv, {t: i})[t]
            assert alone["hash"] == mix[i]["hash"]
            assert alone["residual"] == mix[i]["residual"]
            assert alone["iters"] == mix[i]["iters"]


def test_admission_against_budget(cam):
    # a 1024^2 request needs three arrays of 1028 x 1056 doubles (26 MB);
    # a 40 MiB budget holds one
    p = {"width": 1024, "height": 1024, "iters": 200, "tol": 1e-9, "seed": 3}
    with cam.Server("engine", budget_bytes=40 << 20) as srv:
        t1 = srv.submit(p)
        assert t1 is not None
        assert srv.submit(dict(p, seed=4)) is None  # BUSY: budget in use
        st = srv.stats()
        assert st["active"] == 1 and st["bytes_in_use"] <= st["budget_bytes"]
        first = drain(srv, {t1: 0})[t1]
        t2 = srv.submit(dict(p, seed=4))
        assert t2 is not None
        second = drain(srv, {t2: 0})[t2]
    assert first["iters"] == second["iters"] == 200
    with cam.Jacobi2D(1024, 1024, 200, form="plain", seed=3) as j:
        assert j.solve().hash == first["hash"]


def test_continuous_batching_overlaps_requests(cam):
    # a long request is in flight while short ones arrive, run and finish
    long_p = {"width": 2048, "height": 2048, "iters": 1500, "tol": 1e-12, "seed": 9}
    with cam.Server("engine", budget_bytes=1 << 30) as srv:
        tl = srv.submit(long_p)
        time.sleep(0.01)
        shorts = {srv.submit({"width": 64, "height": 64, "iters": 20,
                              "seed": 50 + i}): i for i in range(8)}
        got = {}
        order = []
        t0 = time.time()
        while len(got) < 9:
            for c in srv.poll():
                got[c["ticket"]] = c
                order.append(c["ticket"])
            assert time.time() - t0 < 60
            time.sleep(0.0005)
    assert order[-1] == tl  # the short ones did not wait for the long one
    for t, i in shorts.items():
        assert got[t]["hash"] == O.grid_hash(O.jacobi_run(64, 64, 20, seed=50 + i)[0])


def test_submit_rejects_bad_request(cam):
    with cam.Server("engine", budget_bytes=64 << 20) as srv:
        with pytest.raises(cam.ChmError):
            srv.submit({"width": 0, "height": 10, "iters": 5})
        with pytest.raises(cam.ChmError):  # larger than the whole budget
            srv.submit({"width": 8192, "height": 8192, "iters": 5})
