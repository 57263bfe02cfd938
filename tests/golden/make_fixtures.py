"""Writes the full-size golden fixtures from the CPU oracle's threaded
message-driven chare runner: per-chare interior hashes and the exact
residual of the last sweep.  Run once on a CPU host (a few minutes):

    python tests/golden/make_fixtures.py
"""
import json
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


def main():
    threads = min(8, os.cpu_count() or 1)
    only = sys.argv[1:]
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
