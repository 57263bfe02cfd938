"""Writes the full-size golden fixtures from the CPU oracle.

  jacobi_4096_8x8_200, jacobi_16384_16x16_1000: the threaded message-driven
      chare runner; per-chare interior hashes and the exact residual of the
      last sweep.
  batch_512: every problem of the 512-problem batch workload
      (charming_amd.workloads.batch_problems), solved monolithically: hash
      of the interior, exact residual of the last sweep, sweeps.
  serve_sample: 16 requests of the serving workload
      (workloads.serve_requests, sides 256..8192, stop at relative residual
      1e-6 or 2000 sweeps): 15 picked at evenly spaced quantiles of the area
      of a 1000-request stream, plus one 8192 x 8192 request; hash, sweeps,
      S_last and S_0 of the oracle's tolerance solve.

Run once on a CPU host (several minutes; the batch and serving cases use a
process per core):

    python tests/golden/make_fixtures.py [name ...]
"""
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

CASES = [
    # name, W, H, chare rows, chare cols, iterations, seed
    ("jacobi_4096_8x8_200", 4096, 4096, 8, 8, 200, 0xC4A40002),
    ("jacobi_16384_16x16_1000", 16384, 16384, 16, 16, 1000, 0xC4A40003),
]


def _batch_one(p):
    g, res = O.jacobi_run(p["width"], p["height"], p["iters"], seed=p["seed"],
                          boundary=p.get("boundary", 1.0))
    return {"hash": str(int(O.grid_hash(g))), "residual": str(res),
            "iters": p["iters"]}


def _serve_one(p):
    g, n, s_last, s0 = O.jacobi_solve_tol(p["width"], p["height"], p["iters"],
                                          p["tol"], seed=p["seed"],
                                          boundary=p.get("boundary", 1.0))
    return {"hash": str(int(O.grid_hash(g))), "iters": int(n),
            "residual": str(s_last), "residual0": str(s0)}


def serve_sample():
    """The 16 requests of the serving fixture (deterministic)."""
    from charming_amd import workloads as WL
    _, reqs = WL.serve_requests(1000, 100.0)
    order = sorted(range(len(reqs)),
                   key=lambda i: (reqs[i]["width"] * reqs[i]["height"], i))
    picks = [order[round(k * (len(order) - 1) / 14)] for k in range(15)]
    sample = [dict(reqs[i], index=i) for i in picks]
    big = dict(reqs[picks[-1]], width=8192, height=8192,
               seed=(WL.SERVE_SEED + 0xB16) & (2**64 - 1), index=-1)
    return sample + [big]


def write_pool_fixture(name, probs, fn, extra):
    t = time.time()
    procs = min(8, os.cpu_count() or 1)
    # largest first, so the pool finishes together
    order = sorted(range(len(probs)), key=lambda i: -probs[i]["width"] *
                   probs[i]["height"] * probs[i]["iters"])
    with mp.get_context("fork").Pool(procs) as pool:
        done = pool.map(fn, [probs[i] for i in order], chunksize=1)
    results = [None] * len(probs)
    for i, r in zip(order, done):
        results[i] = r
    d = dict(extra, problems=probs, results=results)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(d, f, indent=0)
    print(f"{name}: {time.time() - t:.1f} s", flush=True)


def main():
    threads = min(8, os.cpu_count() or 1)
    only = sys.argv[1:]
    if not only or "batch_512" in only:
        from charming_amd import workloads as WL
        write_pool_fixture("batch_512", WL.batch_problems(), _batch_one,
                           {"workload": "charming_amd.workloads."
                                        "batch_problems() (512 problems)"})
    if not only or "serve_sample" in only:
        write_pool_fixture("serve_sample", serve_sample(), _serve_one,
                           {"workload": "16 requests of charming_amd."
                                        "workloads.serve_requests(1000, 100.0)"
                                        " by area quantile, plus 8192^2"})
    for name, W, H, cr, cc, iters, seed in CASES:
        if only and name not in only:
            continue
        t = time.time()
        hashes, res = O.jacobi_chares_hashes(W, H, cr, cc, iters, threads,
                                             seed=seed)
        d = {"W": W, "H": H, "chares": [cr, cc], "iters": iters, "seed": seed,
             "init": "splitmix64 uniform [0,1) interior, boundary 1.0",
             "residual_last_sweep": str(res),
             "chare_hashes": [str(int(h)) for h in hashes],
             "grid_hash": str(int(sum(int(h) for h in hashes) % 2**64))}
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(d, f, indent=0)
        print(f"{name}: {time.time() - t:.1f} s", flush=True)


if __name__ == "__main__":
    main()
