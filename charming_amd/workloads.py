"""Synthetic workloads of the batched and served Jacobi2D forms.

Deterministic (numpy PCG64 seeded by the caller), so a bench run, a test and
the oracle see the same problems.
"""
import numpy as np

BATCH_SEED = 0xC4A40004
SERVE_SEED = 0xC4A40005


def log_uniform_int(rng, lo, hi, n):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), n))).astype(
        np.int64).clip(lo, hi)


def batch_problems(n=512, side_lo=64, side_hi=4096, it_lo=50, it_hi=500,
                   seed=BATCH_SEED):
    """Uneven batch: side s log-uniform in [side_lo, side_hi]; problems with
    even index are s x s squares, odd ones 2:1 rectangles (width s, height
    s/2); iteration counts uniform in [it_lo, it_hi]; seeded uniform [0, 1)
    interiors (seed + index) with boundary 1.0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = log_uniform_int(rng, side_lo, side_hi, n)
    iters = rng.integers(it_lo, it_hi + 1, n)
    probs = []
    for i in range(n):
        s = int(sides[i])
        h = s if i % 2 == 0 else max(1, s // 2)
        probs.append({"width": s, "height": h, "iters": int(iters[i]),
                      "seed": (seed + 1 + i) & (2**64 - 1), "boundary": 1.0})
    return probs


def serve_requests(n, rate, side_lo=256, side_hi=8192, tol=1e-6,
                   max_iters=2000, seed=SERVE_SEED):
    """Poisson stream: arrival times (s, from 0) at `rate` req/s and requests
    whose width and height are each log-uniform in [side_lo, side_hi],
    stopping at relative residual `tol` or `max_iters` sweeps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    ws = log_uniform_int(rng, side_lo, side_hi, n)
    hs = log_uniform_int(rng, side_lo, side_hi, n)
    reqs = []
    for i in range(n):
        reqs.append({"width": int(ws[i]), "height": int(hs[i]),
                     "iters": max_iters, "tol": tol,
                     "seed": (seed + 1 + i) & (2**64 - 1), "boundary": 1.0})
    return arrivals.tolist(), reqs


def cell_updates(probs, iters=None):
    """Interior cell updates of a set of problems (iters overrides)."""
    if iters is None:
        iters = [p["iters"] for p in probs]
    return int(sum(p["width"] * p["height"] * it for p, it in zip(probs, iters)))
