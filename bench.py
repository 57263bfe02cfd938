"""Benchmark of the FP64 Jacobi2D hot path on one MI355X.

Workload (the 16384^2 configuration): a 16384 x 16384 interior, seeded
uniform [0,1) with boundary 1.0, decomposed into 16 x 16 chares of 1024^2
with message-driven halo exchange, 1000 iterations, run in one persistent
launch of the charming_amd scheduler.  One step = one 1000-iteration solve
from u_0 on the grid already resident in HBM; u_0 is re-seeded on the
device between steps, outside the timed region.

Prints ONE JSON line.  `value` is whole-job throughput in Gcell-upd/s
(interior cells x iterations / wall seconds of the timed steps).

Roofline of the dominant kernel (the persistent scheduler launch): the chares
exchange halos D deep and fuse D sweeps per pass over their block (D = 16 by
default for 1024^2 chares), so the algorithm's compulsory HBM bytes are 16/D
per cell-update (read u_k once, write u_{k+D} once per D updates).
`roofline.achieved` is that figure times the launch's cell-updates over its
HIP-event duration; `roofline.traffic` is the bytes the memory system really
moved (rocprof FETCH_SIZE + WRITE_SIZE, calibrated), beside its ratio to the
minimum.  At sixteen fused sweeps the launch is bound by vector-ALU issue
rather than by HBM (profiles/sq_valu_jacobi2d_16384_persistent.md), so the
HBM fraction is not the binding one.  The 16 B-per-update rate of the
unfused sweep is reported separately (`gbs_at_16B_per_update`), and the
one-sweep-per-exchange persistent form, which moves 16 B per update, is
measured beside the plain forms.

Other blocks:
  "compare":  the plain per-launch and hipGraph forms of the same solve, the
              persistent form with one-deep halos, and the streaming copy
              rate of this device;
  "pingpong": two chares on PEs 0 and n/2 exchange seeded payloads of 8 B to
              1 MiB, 1000 round trips each after 10 warm-up, every byte and
              refnum checked; one-way latency and bandwidth per size, and the
              raw mailbox floor without the runtime;
  "batch":    512 uneven problems (sides log-uniform 64..4096, half square and
              half 2:1, 50..500 sweeps) autotuned and solved in one persistent
              engine launch, against the plain per-launch and hipGraph forms;
  "serve":    Poisson request streams (sides log-uniform 256..8192, stop at
              relative residual 1e-6 or 2000 sweeps, 200 GB budget) served by
              the engine (continuous batching) and by the plain per-request
              form: solves/s and p50/p99 latency per arrival rate;
  "cpu_baseline": the oracle's threaded message-driven chare runner on the
              host cores over a bounded sample of the same workload, with the
              512^2 chare runner and single-thread sweeps beside it.
"""
import argparse
import ctypes
import json
import os
import re
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W = H = 16384
CHARES = (16, 16)
ITERS = 1000
SEED = 0xC4A40003
PEAK_HBM_GBS = 8000.0
BYTES_PER_CELL = 16.0  # one unfused sweep: read u once, write u' once (FP64)
MAX_DEPTH = 16         # deepest halo exchange of the persistent form


def _p(q, xs):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(depth, band):
    """HBM bytes per launch of the persistent scheduler, measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and calibrated against a copy
    kernel of known size (profiles/pmc_traffic.json), or None when that
    profile was taken on another workload, halo depth or band height."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        if (d.get("workload") == f"{W}x{H}/{CHARES[0]}x{CHARES[1]}/{ITERS}"
                and d.get("halo_depth") == depth
                and d.get("band_rows") == band):
            return d.get("bytes_per_launch")
    except (OSError, ValueError):
        pass
    return None


def host_cpu():
    """CPU model and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            m = re.search(r"^model name\s*:\s*(.+)$", f.read(), re.M)
            model = m.group(1).strip() if m else None
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    # the host grants a GPU job a CPU share that it exports as the thread
    # count for OpenMP-style pools; without it, every core of the affinity
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return model, aff, (min(aff, share) if share > 0 else aff)


def cpu_baseline(budget_s):
    """The oracle on host cores: the message-driven chare runner over a
    bounded sample of the 16384^2 workload (the headline comparison), plus
    the 512^2 / 4x4-chare / 100-iteration runner and single-thread sweeps
    (median of 5 runs each)."""
    from oracle import oracle as O
    model, aff, cores = host_cpu()
    # 16384^2 sample: scale the iteration count to about budget_s
    O.jacobi_chares_run(W, H, CHARES[0], CHARES[1], 2, cores, seed=SEED,
                        want_grid=False)
    t = O.last_run_seconds()
    iters = max(2, min(ITERS, int(budget_s / max(t / 2, 1e-6))))
    O.jacobi_chares_run(W, H, CHARES[0], CHARES[1], iters, cores, seed=SEED,
                        want_grid=False)
    t = O.last_run_seconds()
    c3 = W * H * iters / t / 1e9

    def med(fn, n=5):
        ts = []
        for _ in range(n):
            ts.append(fn())
        return statistics.median(ts)

    def c1_chares():
        O.jacobi_chares_run(512, 512, 4, 4, 100, min(cores, 16), seed=SEED,
                            want_grid=False)
        return O.last_run_seconds()

    def mono(w, h, it):
        def run():
            t0 = time.perf_counter()
            O.jacobi_run(w, h, it, seed=SEED)
            return time.perf_counter() - t0
        return run
    t_c1 = med(c1_chares)
    t_m1 = med(mono(512, 512, 100))
    t_m2 = med(mono(4096, 4096, 5), 3)

    def rate(cells, t):
        g = cells / t / 1e9
        return {"gcell_per_s": round(g, 3),
                "gbs_at_16B_per_update": round(BYTES_PER_CELL * g, 2)}
    return {"value": round(c3, 3), "unit": "Gcell-upd/s", "cores": cores,
            "kind": "port",
            "sample": f"{W}x{H} interior, {CHARES[0]}x{CHARES[1]} chares, "
                      f"{iters} of the {ITERS} iterations, message-driven "
                      f"chares on {cores} threads; iteration phase timed "
                      f"({t:.1f} s)",
            "cpu_model": model, "affinity_cores": aff,
            "variants": {
                "chares_512x512_4x4_100it": dict(
                    rate(512 * 512 * 100, t_c1), threads=min(cores, 16),
                    median_of=5, seconds=round(t_c1, 4)),
                "single_thread_512x512_100it": dict(
                    rate(512 * 512 * 100, t_m1), threads=1, median_of=5,
                    seconds=round(t_m1, 4)),
                "single_thread_4096x4096_5it": dict(
                    rate(4096 * 4096 * 5, t_m2), threads=1, median_of=3,
                    seconds=round(t_m2, 4)),
            }}


def raw_pingpong(iters):
    """The raw mailbox floor (no runtime), as a child process started before
    this process touches the GPU."""
    exe = os.path.join(ROOT, "charming_amd", "bin", "raw_pingpong")
    if not os.path.exists(exe):
        return None
    p = subprocess.run([exe, str(iters), "10"], capture_output=True,
                       text=True, timeout=120)
    got = re.findall(r"PEs 0 and (\d+) of (\d+), (\d+) round trips: ([\d.]+) "
                     r"us per one-way trip, errors (\d+)", p.stdout)
    return {"rc": p.returncode,
            "placements": [{"pes": [0, int(a)], "n_pes": int(n),
                            "round_trips": int(r), "one_way_us": float(us),
                            "errors": int(e)} for a, n, r, us, e in got]}


def _run_program(prog, args):
    """Runs a chare program library through chm_main in this process and
    returns what it printed (C stdout captured through a temporary file)."""
    lib = os.path.join(ROOT, "charming_amd", "lib")
    rt = ctypes.CDLL(os.path.join(lib, "libcharming_amd.so"),
                     mode=ctypes.RTLD_GLOBAL)
    ctypes.CDLL(os.path.join(lib, f"libchm_{prog}.so"))
    rt.chm_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    rt.chm_select_program.argtypes = [ctypes.c_char_p]
    if rt.chm_select_program(f"{prog}.hip".encode()) != 0:
        raise RuntimeError(f"program {prog} not registered")
    argv = [prog.encode()] + [str(a).encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    libc = ctypes.CDLL(None)
    os.environ["CHM_STATS"] = "1"
    os.environ["CHM_TIMEOUT_S"] = "60"
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    with tempfile.TemporaryFile(mode="w+") as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            rc = rt.chm_main(len(argv), arr)
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        out = tmp.read()
    return rc, out


def bench_pingpong(iters, warmup):
    """C2: payloads 8 B .. 1 MiB, `iters` round trips each after `warmup`,
    timed without the receivers' byte check (which would be most of the time
    at large sizes); a second run of 20 round trips per size checks every
    byte."""
    rc, out = _run_program("pingpong", [8, 1 << 20, iters, warmup, 0])
    rc2, out2 = _run_program("pingpong", [8, 1 << 20, 20, 2, 1])
    sizes = re.findall(r"Size (\d+) took ([\d.]+) us", out)
    m = re.search(r"round trips (\d+), payload errors (\d+)", out2)
    st = dict(re.findall(r"(\w+)=([\d.]+)", out.split("[STATS]")[-1]))
    rows = [{"bytes": int(s), "one_way_us": float(us),
             "gb_per_s": round(int(s) / (float(us) * 1e-6) / 1e9, 3)}
            for s, us in sizes]
    m1 = re.search(r"round trips (\d+), payload errors", out)
    return {"workload": f"2 chares on PEs 0 and n/2, payloads 8 B .. 1 MiB "
                        f"doubling, {iters} round trips each after {warmup} "
                        "warm-up, every reply matched by refnum; seeded "
                        "bytes checked in a second run of 20 round trips "
                        "per size",
            "rc": rc, "sizes": rows,
            "round_trips": int(m1.group(1)) if m1 else None,
            "checked_run": {"rc": rc2,
                            "round_trips": int(m.group(1)) if m else None,
                            "payload_errors": int(m.group(2)) if m else None},
            "deferred": int(st.get("deferred", -1)),
            "n_pes": int(st.get("pes", 0))}


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
    best = min(eng, eng2, key=lambda o: o["result"].device_ms)
    e = rate_of(best)
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
                                  best["result"].device_ms, 2),
        "speedup_vs_graph": round(graph["result"].device_ms /
                                  best["result"].device_ms, 2),
        "autotune": {"config": kw, "tiles": bcount, "sample_iters": sample_iters,
                     "sample_gcell_per_s": round(rate / 1e9, 2),
                     "seconds": round(tune_s, 1)},
        "bit_identical_across_forms": same,
    }


def bench_serve(cam, plan, budget):
    """Poisson request streams, engine (continuous batching) vs plain; plan =
    [(rate, requests)]."""
    from charming_amd import serve as SV
    from charming_amd import workloads as WL
    out = {"workload": "Poisson arrivals; sides log-uniform 256..8192 per "
                       "side, stop at relative residual 1e-6 or 2000 sweeps; "
                       "latency = arrival to completion seen by the request "
                       "loop (admission waits included)",
           "budget_bytes": budget, "rates": []}
    same = True
    for rate, n_req in plan:
        arr, reqs = WL.serve_requests(n_req, rate)
        row = {"offered_rate": rate, "n": n_req}
        res = {}
        for mode in ("engine", "plain"):
            r = SV.run_stream(mode, arr, reqs, budget_bytes=budget)
            res[mode] = r["results"]
            row[mode] = r["summary"]
            log(json.dumps({"serve_rate": rate, "n": n_req, mode: r["summary"]}))
        same = same and all(a["hash"] == b["hash"] and a["iters"] == b["iters"]
                            for a, b in zip(res["engine"], res["plain"]))
        row["solves_ratio"] = round(row["engine"]["solves_per_s"] /
                                    row["plain"]["solves_per_s"], 2)
        out["rates"].append(row)
    out["bit_identical_engine_vs_plain"] = same
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-pingpong", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--tune-iters", type=int, default=0,
                    help="sweeps per problem in the autotuner's samples "
                         "(0 = every problem's own count)")
    ap.add_argument("--serve-plan", default="10:100,50:250,100:1000,500:300,"
                    "2000:1000", help="rate:requests,... of the serving block")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--pingpong-iters", type=int, default=1000)
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

    # the raw floor runs as a child process, before this one uses the GPU
    raw = None
    if not args.no_pingpong and not args.calibrate:
        raw = raw_pingpong(args.pingpong_iters)

    import charming_amd as cam
    L = cam.jacobi_lib()
    L.chm_hbm_copy_probe.argtypes = [ctypes.c_size_t, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double)]
    if args.calibrate:
        ms = ctypes.c_double()
        rc = L.chm_hbm_copy_probe(1 << 30, 3, ctypes.byref(ms))
        print(json.dumps({"calibration_copy_bytes": 1 << 30,
                          "launches": 8, "ms_per_copy": ms.value, "rc": rc}),
              flush=True)
        return

    log("persistent 16384^2 solve")
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
    # halo depth actually run: the default is as deep as the chares allow
    side = min(W // CHARES[1], H // CHARES[0])
    if args.depth:
        depth = args.depth
    elif nc <= 1 and u <= 1:
        depth = min(side, MAX_DEPTH)
    else:
        depth = 2
    min_bytes = BYTES_PER_CELL / depth * cells
    # counters exist for one configuration (the profile names it)
    traffic = traffic_from_profiles(depth, args.band) if not (nc or u) else None
    kavg = statistics.mean(kern_ms)
    achieved = min_bytes / (kavg / 1e3) / 1e9

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
                   "halo_depth": depth, "sweep_shape": args.shape,
                   "band_rows": args.band or "default"},
        # the unfused sweep's 16 B per cell-update at this solve rate (above
        # the HBM peak once sweeps are fused: it is not a bandwidth)
        "gbs_at_16B_per_update": round(BYTES_PER_CELL * value, 1),
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
                     "traffic_ratio": (round(traffic / min_bytes, 3)
                                       if traffic else None),
                     "bytes_per_cell_update": BYTES_PER_CELL / depth,
                     "algorithmic_bytes_per_launch": min_bytes,
                     "kernel": "charm::rt::scheduler_kernel<512>",
                     "kernel_ms": round(kavg, 3)},
    }

    if not args.no_compare:
        log("compare: copy probe, plain, graph, persistent depth 1")
        ms = ctypes.c_double()
        if L.chm_hbm_copy_probe(4 << 30, 5, ctypes.byref(ms)) == 0:
            out["copy_gbs"] = round(2 * (4 << 30) / ms.value / 1e6, 1)
        comp = {}
        for form, dep in (("plain", 0), ("graph", 0), ("persistent", 1)):
            p = cam.Jacobi2D(W, H, ITERS, form=form, seed=SEED,
                             chares=CHARES if form == "persistent" else None)
            if dep:
                p.set_tuning(depth=dep)
            p.solve()
            p.reset()
            lat, kms = [], []
            for _ in range(max(2, args.steps // 2)):
                s0 = time.perf_counter()
                r = p.solve()
                lat.append((time.perf_counter() - s0) * 1e3)
                kms.append(r.device_ms)
                p.reset()
            p.close()
            g = cells / (statistics.median(lat) / 1e3) / 1e9
            kg = BYTES_PER_CELL * cells / (statistics.mean(kms) / 1e3) / 1e9
            name = form if not dep else "persistent_depth1"
            comp[name] = {"gcell_per_s": round(g, 2),
                          "gbs_at_16B_per_update": round(BYTES_PER_CELL * g, 1),
                          "kernel_gbs": round(kg, 1),
                          "kernel_frac_of_peak": round(kg / PEAK_HBM_GBS, 4),
                          "latency_ms": {"p50": round(_p(50, lat), 3),
                                         "p99": round(_p(99, lat), 3)},
                          "launches_per_solve": r.launches}
        out["compare"] = comp

    if not args.no_pingpong:
        log("pingpong")
        pp = bench_pingpong(args.pingpong_iters, 10)
        pp["raw_floor"] = raw
        out["pingpong"] = pp
    if not args.no_batch:
        log("batch")
        out["batch"] = bench_batch(cam, args.tune_iters)
    if not args.no_serve:
        plan = [(float(a), int(b)) for a, b in
                (x.split(":") for x in args.serve_plan.split(",") if x)]
        out["serve"] = bench_serve(cam, plan,
                                   int(args.serve_budget_gb * 1e9))
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
