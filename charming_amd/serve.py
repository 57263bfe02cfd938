"""Host request loop of the served Jacobi2D form.

Replays an arrival schedule against a charming_amd.Server: each request is
offered at its arrival time, admitted in FIFO order as soon as the server
accepts it (admission waits while the HBM budget or the resident-request
limit is exhausted), and its completion is collected by polling.  Latency is
arrival to completion seen by this loop, so it includes admission waits.
"""
import collections
import time

import charming_amd as cam


def _pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def run_stream(mode, arrivals, reqs, budget_bytes=16 << 30, timeout_s=600.0,
               **server_kw):
    """Serve reqs arriving at `arrivals` (seconds from start) on a fresh
    server in `mode` ('engine' or 'plain').  Returns a dict of per-request
    results (by index) and the latency / throughput summary."""
    n = len(reqs)
    by_ticket = {}
    results = [None] * n
    backlog = collections.deque()
    nxt = 0
    done = 0
    max_resident = 0
    with cam.Server(mode, budget_bytes=budget_bytes, **server_kw) as srv:
        t0 = time.perf_counter()
        while done < n:
            now = time.perf_counter() - t0
            if now > timeout_s:
                raise TimeoutError(f"{mode}: {done}/{n} requests after {now:.0f} s")
            busy = False
            while nxt < n and arrivals[nxt] <= now:
                backlog.append(nxt)
                nxt += 1
            while backlog:
                i = backlog[0]
                t = srv.submit(reqs[i])
                if t is None:
                    break
                by_ticket[t] = i
                backlog.popleft()
                busy = True
            max_resident = max(max_resident, len(by_ticket) - done)
            for c in srv.poll():
                i = by_ticket[c["ticket"]]
                c["t_done"] = time.perf_counter() - t0
                c["latency_s"] = c["t_done"] - arrivals[i]
                results[i] = c
                done += 1
                busy = True
            if not busy:
                time.sleep(0.0002)
    lat = [r["latency_s"] * 1e3 for r in results]
    span = max(r["t_done"] for r in results) - arrivals[0]
    cells = sum(p["width"] * p["height"] * r["iters"] for p, r in zip(reqs, results))
    return {
        "results": results,
        "summary": {
            "mode": mode, "requests": n,
            "offered_rate": round(n / max(arrivals[-1] - arrivals[0], 1e-9), 2)
            if n > 1 else None,
            "solves_per_s": round(n / span, 3),
            "latency_ms": {"p50": round(_pct(lat, 50), 2),
                           "p99": round(_pct(lat, 99), 2),
                           "max": round(max(lat), 2)},
            "gcell_per_s": round(cells / span / 1e9, 2),
            "max_resident": max_resident,
        },
    }
