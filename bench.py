"""Benchmark of the FP64 Jacobi2D hot path on one MI355X.

Workload (the 16384^2 configuration): a 16384 x 16384 interior, seeded
uniform [0,1) with boundary 1.0, decomposed into 16 x 16 chares of 1024^2
with message-driven halo exchange, 1000 iterations, run in one persistent
launch of the charming_amd scheduler.  One step = one 1000-iteration solve
on the grid already resident in HBM (each step continues from the previous
state; 1000 is even, so the buffer parity returns to the start).

Prints ONE JSON line.  `value` is whole-job throughput in Gcell-upd/s
(interior cells x iterations / wall seconds of the timed steps).  Beside it:
the plain per-launch and hipGraph forms of the same sweep, the measured copy
ceiling, the roofline of the dominant kernel (16 algorithmic bytes per
cell-update over its HIP-event duration) and the CPU baseline (the oracle's
threaded message-driven runner on the host cores, bounded sample).

Two more blocks, each bit-checked across forms:
  "batch": 512 uneven problems (sides log-uniform 64..4096, half square and
           half 2:1, 50..500 sweeps) autotuned and solved in one persistent
           engine launch, against the same problems one after another in the
           plain per-launch form and as one hipGraph per problem;
  "serve": Poisson request streams (sides log-uniform 256..8192, stop at
           relative residual 1e-6 or 2000 sweeps, 200 GB budget) served by
           the engine (continuous batching) and by the plain per-request
           form: solves/s and p50/p99 latency per arrival rate.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W = H = 16384
CHARES = (16, 16)
ITERS = 1000
SEED = 0xC4A40003
PEAK_HBM_GBS = 8000.0
BYTES_PER_CELL = 16.0  # read u once, write u' once (FP64)


def _p(q, xs):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def traffic_from_profiles():
    """HBM bytes per launch of the persistent scheduler, measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and calibrated against a copy
    kernel of known size (profiles/pmc_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        if d.get("workload") == f"{W}x{H}/{CHARES[0]}x{CHARES[1]}/{ITERS}":
            return d.get("bytes_per_launch")
    except (OSError, ValueError):
        pass
    return None


def cpu_baseline(budget_s):
    """The oracle's message-driven chare runner on host threads, over a
    bounded sample of the same workload (same grid and chares, fewer
    iterations)."""
    from oracle import oracle as O
    cores = min(16, len(os.sched_getaffinity(0)))
    iters = 2
    O.jacobi_chares_run(W, H, CHARES[0], CHARES[1], iters, cores, seed=SEED,
                        want_grid=False)
    t = O.last_run_seconds()
    # scale the iteration count so the timed sample is about budget_s
    iters = max(2, min(ITERS, int(budget_s / max(t / 2, 1e-6))))
    O.jacobi_chares_run(W, H, CHARES[0], CHARES[1], iters, cores, seed=SEED,
                        want_grid=False)
    t = O.last_run_seconds()
    return {"value": W * H * iters / t / 1e9, "unit": "Gcell-upd/s",
            "cores": cores, "kind": "port",
            "sample": f"{W}x{H} interior, {CHARES[0]}x{CHARES[1]} chares, "
                      f"{iters} of the {ITERS} iterations, message-driven "
                      f"threads; iteration phase timed ({t:.1f} s)"}


def bench_batch(cam, sample_iters):
    """512 uneven problems: autotuned engine launch vs plain and graph."""
    from charming_amd import workloads as WL
    probs = WL.batch_problems()
    t0 = time.perf_counter()
    cfg, blocks, rate = cam.batch_autotune(probs, sample_iters=sample_iters)
    tune_s = time.perf_counter() - t0
    tuned = [dict(p, block=b) for p, b in zip(probs, blocks)]
    kw = {k: cfg[k] for k in ("wg_threads", "panel_chunks", "rows_in_flight")}
    eng = cam.batch_run(tuned, **kw)
    eng2 = cam.batch_run(tuned, **kw)
    plain = cam.batch_plain(probs)
    graph = cam.batch_plain(probs, use_graph=True)
    same = (eng["hashes"] == plain["hashes"] == graph["hashes"] == eng2["hashes"]
            and eng["residuals"] == plain["residuals"] == graph["residuals"])
    cells = WL.cell_updates(probs)

    def rate_of(o):
        ms = o["result"].device_ms
        return {"ms": round(ms, 2), "gcell_per_s": round(cells / ms / 1e6, 2),
                "solves_per_s": round(len(probs) / ms * 1e3, 1)}
    e = rate_of(min(eng, eng2, key=lambda o: o["result"].device_ms))
    r = eng["result"]
    bcount = {b: blocks.count(b) for b in sorted(set(blocks))}
    return {
        "workload": "512 problems, sides log-uniform 64..4096 (even index "
                    "square, odd 2:1), sweeps uniform 50..500, seeded",
        "cell_updates": cells,
        "engine": dict(e, tasks=r.tasks, steals=r.steals, n_wg=r.n_wg,
                       wg_threads=r.wg_threads),
        "plain": rate_of(plain), "graph": rate_of(graph),
        "speedup_vs_plain": round(plain["result"].device_ms /
                                  min(eng["result"].device_ms,
                                      eng2["result"].device_ms), 2),
        "speedup_vs_graph": round(graph["result"].device_ms /
                                  min(eng["result"].device_ms,
                                      eng2["result"].device_ms), 2),
        "autotune": {"config": kw, "tiles": bcount, "sample_iters": sample_iters,
                     "sample_gcell_per_s": round(rate / 1e9, 2),
                     "seconds": round(tune_s, 1)},
        "bit_identical_across_forms": same,
    }


def bench_serve(cam, rates, n_req, budget):
    """Poisson request streams, engine (continuous batching) vs plain."""
    from charming_amd import serve as SV
    from charming_amd import workloads as WL
    out = {"workload": f"{n_req} requests per rate, sides log-uniform "
                       "256..8192 per side, stop at relative residual 1e-6 "
                       "or 2000 sweeps",
           "budget_bytes": budget, "rates": []}
    same = True
    for rate in rates:
        arr, reqs = WL.serve_requests(n_req, rate)
        row = {"offered_rate": rate}
        res = {}
        for mode in ("engine", "plain"):
            r = SV.run_stream(mode, arr, reqs, budget_bytes=budget)
            res[mode] = r["results"]
            row[mode] = r["summary"]
        same = same and all(a["hash"] == b["hash"] and a["iters"] == b["iters"]
                            for a, b in zip(res["engine"], res["plain"]))
        row["solves_ratio"] = round(row["engine"]["solves_per_s"] /
                                    row["plain"]["solves_per_s"], 2)
        out["rates"].append(row)
        print(json.dumps({"serve_rate": rate, "engine": row["engine"],
                          "plain": row["plain"]}), file=sys.stderr, flush=True)
    out["bit_identical_engine_vs_plain"] = same
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--tune-iters", type=int, default=20,
                    help="sweeps per problem in the autotuner's samples")
    ap.add_argument("--serve-rates", default="10,100,2000")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--depth", type=int, default=0,
                    help="halo depth of the persistent form = sweeps fused "
                         "per exchange (0 = the library default)")
    ap.add_argument("--shape", default="0,0",
                    help="fused-sweep shape: panel chunks,rows in flight "
                         "(0,0 = the library default)")
    ap.add_argument("--band", type=int, default=0,
                    help="rows per band of the deep fused sweep (0 = the "
                         "library default)")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the streaming copy of a known size (1 GiB "
                         "read + 1 GiB written per launch) used to scale "
                         "traffic counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not "
              "supported (this build runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import charming_amd as cam
    L = cam.jacobi_lib()
    if args.calibrate:
        import ctypes
        ms = ctypes.c_double()
        L.chm_hbm_copy_probe.argtypes = [ctypes.c_size_t, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]
        rc = L.chm_hbm_copy_probe(1 << 30, 3, ctypes.byref(ms))
        print(json.dumps({"calibration_copy_bytes": 1 << 30, "launches": 4,
                          "ms_per_copy": ms.value, "rc": rc}), flush=True)
        return

    j = cam.Jacobi2D(W, H, ITERS, form="persistent", seed=SEED, chares=CHARES)
    nc, u = (int(x) for x in args.shape.split(","))
    if args.depth or nc or u or args.band:
        j.set_tuning(nc, u, band_rows=args.band, depth=args.depth)
    for _ in range(args.warmup):
        j.solve()
        j.reset()
    kern_ms, step_ms = [], []
    msgs = deferred = 0
    # each step solves from u_0: the re-seeding between steps (reset) is
    # outside the timed region
    for _ in range(args.steps):
        s0 = time.perf_counter()
        r = j.solve()
        step_ms.append((time.perf_counter() - s0) * 1e3)
        kern_ms.append(r.device_ms)
        msgs, deferred = r.messages, r.deferred
        j.reset()
    wall = sum(step_ms) / 1e3
    j.close()

    cells = float(W) * H * ITERS
    value = cells * args.steps / wall / 1e9
    # the counters were taken on the default configuration only
    default_cfg = not (args.depth or args.band or nc or u)
    traffic = traffic_from_profiles() if default_cfg else None
    kavg = statistics.mean(kern_ms)
    achieved = BYTES_PER_CELL * cells / (kavg / 1e3) / 1e9

    out = {
        "metric": "Jacobi2D Gcell-upd/s & HBM GB/s; p50/p99 solve latency vs "
                  "plain, 1 GPU",
        "value": round(value, 3),
        "unit": "Gcell-upd/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (splitmix64 uniform [0,1) interior, boundary 1.0)",
        "config": {"workload": f"jacobi2d {W}x{H}, {CHARES[0]}x{CHARES[1]} "
                               f"chares, {ITERS} iterations, persistent "
                               "message-driven",
                   "grid": [W, H], "chares": list(CHARES), "iters": ITERS,
                   "form": "persistent", "parallelism": "1 GPU",
                   "halo_depth": args.depth or "default",
                   "sweep_shape": args.shape,
                   "band_rows": args.band or "default"},
        "hbm_gbs": round(BYTES_PER_CELL * value, 1),
        # `achieved` counts the algorithmic 16 B per cell-update; the chares
        # sweep several times per pass (deep halos), so the bytes really moved
        # (rocprof counters, `traffic`) are fewer: their rate over the same
        # launch time is the memory system's actual load
        "hbm_traffic_gbs": (round(traffic / (kavg / 1e3) / 1e9, 1)
                            if traffic else None),
        "latency_ms": {"p50": round(_p(50, step_ms), 3),
                       "p99": round(_p(99, step_ms), 3)},
        "messages_per_solve": msgs,
        "deferred_per_solve": deferred,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic,
                     "kernel": "charm::rt::scheduler_kernel<1024>",
                     "kernel_ms": round(kavg, 3)},
    }

    if not args.no_compare:
        import ctypes
        ms = ctypes.c_double()
        L.chm_hbm_copy_probe.argtypes = [ctypes.c_size_t, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]
        if L.chm_hbm_copy_probe(4 << 30, 5, ctypes.byref(ms)) == 0:
            out["copy_ceiling_gbs"] = round(2 * (4 << 30) / ms.value / 1e6, 1)
        comp = {}
        for form in ("plain", "graph"):
            p = cam.Jacobi2D(W, H, ITERS, form=form, seed=SEED)
            p.solve()
            p.reset()
            lat = []
            for _ in range(max(2, args.steps // 2)):
                s0 = time.perf_counter()
                r = p.solve()
                lat.append((time.perf_counter() - s0) * 1e3)
                p.reset()
            p.close()
            g = cells / (statistics.median(lat) / 1e3) / 1e9
            comp[form] = {"gcell_per_s": round(g, 2),
                          "hbm_gbs": round(BYTES_PER_CELL * g, 1),
                          "latency_ms": {"p50": round(_p(50, lat), 3),
                                         "p99": round(_p(99, lat), 3)},
                          "launches_per_solve": ITERS}
        out["compare"] = comp

    if not args.no_batch:
        out["batch"] = bench_batch(cam, args.tune_iters)
    if not args.no_serve:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        out["serve"] = bench_serve(cam, rates, args.serve_requests,
                                   int(args.serve_budget_gb * 1e9))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
