"""The batched and served configurations at their full size on the GPU,
against committed oracle fixtures (tests/golden/make_fixtures.py).

batch_512: the 512-problem workload (sides log-uniform 64..4096, half 2:1
rectangles, 50..500 sweeps) autotuned and solved in one engine launch, and
the same problems in the plain per-launch form and as one hipGraph per
problem; every problem's hash, exact residual and sweep count must equal the
oracle's in all three.

serve_sample: 16 serving requests (sides 256..8192 including 8192 x 8192,
stop at relative residual 1e-6 or 2000 sweeps) submitted together to an
engine server with the 200 GB budget and the default load-sized tiles, and
again one at a time; each result must equal the oracle's tolerance solve.
"""
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(name):
    with open(os.path.join(HERE, name + ".json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def cam():
    import charming_amd
    charming_amd.jacobi_lib()
    return charming_amd


def _check_batch(out, fx, form):
    bad = []
    for i, r in enumerate(fx["results"]):
        got = (out["hashes"][i], out["residuals"][i], out["iters"][i])
        want = (int(r["hash"]), int(r["residual"]), r["iters"])
        if got != want:
            bad.append((i, got, want))
    assert not bad, f"{form}: {len(bad)} problems differ, first {bad[:3]}"


def test_c4_full_batch_matches_fixture(cam):
    fx = _fixture("batch_512")
    from charming_amd import workloads as WL
    probs = WL.batch_problems()
    assert [(p["width"], p["height"], p["iters"], p["seed"]) for p in probs] == \
        [(p["width"], p["height"], p["iters"], p["seed"]) for p in fx["problems"]]
    print("autotuning the 512-problem batch", flush=True)
    cfg, blocks, _ = cam.batch_autotune(probs, sample_iters=10)
    tuned = [dict(p, block=b) for p, b in zip(probs, blocks)]
    kw = {k: cfg[k] for k in ("wg_threads", "panel_chunks", "rows_in_flight")}
    print(f"engine launch, config {kw}", flush=True)
    _check_batch(cam.batch_run(tuned, **kw), fx, "engine (autotuned)")
    print("engine launch, default tiles", flush=True)
    _check_batch(cam.batch_run(probs), fx, "engine (default)")
    print("plain per-launch form", flush=True)
    _check_batch(cam.batch_plain(probs), fx, "plain")
    print("hipGraph form", flush=True)
    _check_batch(cam.batch_plain(probs, use_graph=True), fx, "graph")


def _serve_all(cam, reqs, budget, one_at_a_time=False):
    import time
    results = {}
    with cam.Server("engine", budget_bytes=budget) as srv:
        pending = list(range(len(reqs)))
        tickets = {}
        t0 = time.time()
        while len(results) < len(reqs):
            assert time.time() - t0 < 90, f"{len(results)} of {len(reqs)} done"
            while pending and (not one_at_a_time or not tickets):
                t = srv.submit(reqs[pending[0]])
                if t is None:
                    break
                tickets[t] = pending.pop(0)
            for c in srv.poll():
                results[tickets.pop(c["ticket"])] = c
            time.sleep(0.001)
    return [results[i] for i in range(len(reqs))]


@pytest.mark.parametrize("together", [True, False])
def test_c5_sample_matches_fixture(cam, together):
    fx = _fixture("serve_sample")
    reqs = [{k: p[k] for k in ("width", "height", "iters", "tol", "seed",
                               "boundary")} for p in fx["problems"]]
    print(f"serving {len(reqs)} requests, together={together}", flush=True)
    got = _serve_all(cam, reqs, 200 * 10**9, one_at_a_time=not together)
    bad = []
    for i, (g, r) in enumerate(zip(got, fx["results"])):
        want = (int(r["hash"]), r["iters"], int(r["residual"]),
                int(r["residual0"]))
        have = (g["hash"], g["iters"], g["residual"], g["residual0"])
        if have != want:
            bad.append((i, reqs[i]["width"], reqs[i]["height"], have, want))
    assert not bad, f"{len(bad)} requests differ, first {bad[:3]}"
