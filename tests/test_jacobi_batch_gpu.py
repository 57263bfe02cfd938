"""Batched form: many uneven Jacobi2D problems in one persistent engine
launch.  Every problem's result (hash of the interior, exact residual of the
last sweep, sweeps done) must equal the monolithic CPU oracle's, whatever
the tiling, workgroup size, sweep shape or company it ran in."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cam():
    import charming_amd
    charming_amd.jacobi_lib()
    return charming_amd


def mixed_problems(n, seed, side_hi=300, it_hi=40):
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(n):
        probs.append({"width": int(rng.integers(1, side_hi)),
                      "height": int(rng.integers(1, side_hi)),
                      "iters": int(rng.integers(1, it_hi)),
                      "block": int(rng.choice([16, 32, 48, 64, 128, 0])),
                      "seed": seed * 1000 + i,
                      "init_mode": int(i % 5 == 4),
                      "boundary": 1.0})
    return probs


def oracle_fixed(p):
    g, res = O.jacobi_run(p["width"], p["height"], p["iters"], seed=p["seed"],
                          mode=p.get("init_mode", 0),
                          boundary=p.get("boundary", 1.0))
    return O.grid_hash(g), res


def check_against_oracle(probs, out):
    for i, p in enumerate(probs):
        h, res = oracle_fixed(p)
        assert out["iters"][i] == p["iters"], (i, p)
        assert out["hashes"][i] == h, (i, p)
        assert out["residuals"][i] == res, (i, p)


def test_mixed_batch_matches_oracle(cam):
    probs = mixed_problems(40, 1)
    print("batch of", len(probs), flush=True)
    out = cam.batch_run(probs)
    check_against_oracle(probs, out)
    r = out["result"]
    assert r.cell_updates == sum(p["width"] * p["height"] * p["iters"] for p in probs)


@pytest.mark.parametrize("threads,nc,u", [(256, 1, 2), (512, 1, 4), (1024, 2, 2),
                                          (256, 2, 2)])
def test_engine_shapes_bit_identical(cam, threads, nc, u):
    probs = mixed_problems(24, 2, side_hi=600, it_hi=20)
    print("shape", threads, nc, u, flush=True)
    out = cam.batch_run(probs, wg_threads=threads, panel_chunks=nc,
                        rows_in_flight=u)
    check_against_oracle(probs, out)


@pytest.mark.parametrize("threads", [256, 512])
def test_sixteen_deep_rounds_match_oracle(cam, threads):
    # engines of up to 512-thread workgroups sweep fixed-count problems in
    # rounds of sixteen: several full rounds, last rounds of 1..15 sweeps
    # (two-sweep and copying passes), tiles of 16..384 and thinner grids
    # that fall back to eight or two
    rng = np.random.default_rng(40 + threads)
    probs = []
    for i in range(20):
        probs.append({"width": int(rng.integers(12, 420)),
                      "height": int(rng.integers(12, 420)),
                      "iters": int(rng.integers(16, 90)),
                      "block": int(rng.choice([16, 40, 96, 192, 384, 0])),
                      "seed": 4000 + threads + i, "init_mode": int(i % 4 == 3),
                      "boundary": 1.0})
    probs += [{"width": 300, "height": 200, "iters": k, "seed": 4500 + k}
              for k in (31, 41, 47, 48)]
    print("sixteen-deep batch", threads, len(probs), flush=True)
    out = cam.batch_run(probs, wg_threads=threads)
    check_against_oracle(probs, out)


def test_problem_alone_equals_in_mix(cam):
    probs = mixed_problems(30, 3, side_hi=400, it_hi=30)
    mix = cam.batch_run(probs)
    for i in (0, 7, 19):
        alone = cam.batch_run([probs[i]])
        assert alone["hashes"][0] == mix["hashes"][i]
        assert alone["residuals"][0] == mix["residuals"][i]


def test_tiles_do_not_change_results(cam):
    base = {"width": 700, "height": 450, "iters": 13, "seed": 77}
    outs = [cam.batch_run([dict(base, block=b)])
            for b in (16, 64, 112, 128, 224, 256, 448, 512, 1024)]
    h, res = oracle_fixed(base)
    for o in outs:
        assert o["hashes"][0] == h and o["residuals"][0] == res


def test_single_tile_and_thin_problems(cam):
    probs = [{"width": 1, "height": 1, "iters": 3, "seed": 1},
             {"width": 1, "height": 500, "iters": 4, "seed": 2, "block": 64},
             {"width": 500, "height": 1, "iters": 5, "seed": 3, "block": 64},
             {"width": 130, "height": 2, "iters": 1, "seed": 4, "block": 16},
             {"width": 64, "height": 64, "iters": 1, "seed": 5}]
    out = cam.batch_run(probs)
    check_against_oracle(probs, out)


def test_tolerance_stop_matches_oracle(cam):
    probs = []
    for i, (w, h, tol) in enumerate([(40, 30, 1e-2), (100, 60, 3e-2),
                                     (33, 77, 5e-3), (64, 64, 0.5),
                                     (200, 150, 2e-2), (17, 5, 1e-3)]):
        probs.append({"width": w, "height": h, "iters": 400, "tol": tol,
                      "seed": 900 + i, "block": [16, 32, 64][i % 3]})
    probs.append({"width": 50, "height": 50, "iters": 6, "tol": 1e-9,
                  "seed": 5})  # the cap binds
    out = cam.batch_run(probs)
    for i, p in enumerate(probs):
        g, n, s_last, s0 = O.jacobi_solve_tol(p["width"], p["height"], p["iters"],
                                              p["tol"], seed=p["seed"])
        assert out["iters"][i] == n, (i, p)
        assert out["hashes"][i] == O.grid_hash(g), (i, p)
        assert out["residuals"][i] == s_last, (i, p)
    assert out["iters"][-1] == 6
    assert any(0 < it < 400 for it in out["iters"][:-1])


def test_tolerance_stop_at_every_sweep_of_a_round(cam):
    # tolerance rounds fuse six sweeps (an FP64 filter in the rounds between
    # the first and the last, exact residuals where it cannot rule a stop
    # out); a stop on sweep 1..5 of a round is repaired by redoing that many
    # sweeps from the round's input, a stop on sweep 6 keeps the round's
    # output; caps that leave a short last round (7, 9, 10 sweeps) too
    tols = [0.5 * 0.7 ** k for k in range(16)]
    probs = [{"width": 48, "height": 40, "iters": 400, "tol": t, "seed": 31,
              "block": 16} for t in tols]
    probs += [{"width": 40, "height": 44, "iters": it, "tol": 1e-12, "seed": it,
               "block": 16} for it in (7, 9, 10)]
    # small grids that converge below the FP64 filter's bound: the rounds
    # it cannot rule out are recomputed exactly, mostly without a stop,
    # until the exact residual reaches 0 (8x8, 10x8) or the cap (8x12)
    probs += [{"width": w, "height": h, "iters": 400, "tol": 1e-13,
               "seed": w * h, "block": 16} for w, h in ((8, 8), (8, 12), (10, 8))]
    out = cam.batch_run(probs)
    stops = set()
    for i, p in enumerate(probs):
        g, n, s_last, s0 = O.jacobi_solve_tol(p["width"], p["height"], p["iters"],
                                              p["tol"], seed=p["seed"])
        assert out["iters"][i] == n, (i, p)
        assert out["hashes"][i] == O.grid_hash(g), (i, p)
        assert out["residuals"][i] == s_last, (i, p)
        stops.add(n % 6)
    assert stops == {0, 1, 2, 3, 4, 5}
    assert out["iters"][-3:] == [328, 400, 387]


def test_plain_and_graph_forms_match_engine(cam):
    probs = mixed_problems(12, 4, side_hi=500, it_hi=25)
    for p in probs:
        p["block"] = 0
    eng = cam.batch_run(probs)
    plain = cam.batch_plain(probs)
    graph = cam.batch_plain(probs, use_graph=True)
    assert plain["hashes"] == eng["hashes"] == graph["hashes"]
    assert plain["residuals"] == eng["residuals"] == graph["residuals"]
    check_against_oracle(probs, eng)


def test_batch_workload_sample_matches_oracle(cam):
    # the uneven batch workload at reduced count and sweeps: sides 64..4096
    from charming_amd import workloads as WL
    probs = WL.batch_problems(n=24, side_hi=1024, it_lo=5, it_hi=12)
    out = cam.batch_run(probs)
    check_against_oracle(probs, out)


def test_large_problem_in_engine_matches_persistent_and_plain(cam):
    # 4096^2 with the default 448^2 tiles (the last row and column of tiles
    # 64 wide) next to small problems; compare with the plain form's device
    # hash (the oracle would take minutes here)
    big = {"width": 4096, "height": 4096, "iters": 30, "seed": 11}
    small = mixed_problems(20, 5, side_hi=200, it_hi=30)
    out = cam.batch_run([big] + small)
    with cam.Jacobi2D(4096, 4096, 30, form="plain", seed=11) as j:
        r = j.solve()
    assert out["hashes"][0] == r.hash
    assert out["residuals"][0] == r.residual
    check_against_oracle(small, {k: v[1:] for k, v in out.items() if k != "result"})


def test_autotune_small_batch(cam):
    probs = mixed_problems(16, 6, side_hi=900, it_hi=10)
    for p in probs:
        p["block"] = 0
    cfg, blocks, rate = cam.batch_autotune(probs, sample_iters=5)
    assert cfg["wg_threads"] in (256, 512, 1024)
    assert (cfg["panel_chunks"], cfg["rows_in_flight"]) in ((1, 2), (1, 4), (2, 2))
    # the powers of two and the multiples of the deep pass's window: 112
    # columns at eight sweeps per round (1024-thread workgroups), 96 at
    # sixteen (512 threads or fewer)
    if cfg["wg_threads"] <= 512:
        assert all(b in (64, 96, 128, 192, 256, 384, 512) for b in blocks)
    else:
        assert all(b in (64, 112, 128, 224, 256, 448, 512) for b in blocks)
    assert rate > 0
    tuned = [dict(p, block=b) for p, b in zip(probs, blocks)]
    out = cam.batch_run(tuned, wg_threads=cfg["wg_threads"],
                        panel_chunks=cfg["panel_chunks"],
                        rows_in_flight=cfg["rows_in_flight"])
    check_against_oracle(probs, out)


def test_bad_arguments_raise(cam):
    with pytest.raises(cam.ChmError):
        cam.batch_run([{"width": 0, "height": 5, "iters": 1}])
    with pytest.raises(cam.ChmError):
        cam.batch_run([{"width": 5, "height": 5, "iters": 1}], wg_threads=128)
    with pytest.raises(cam.ChmError):
        cam.batch_run([{"width": 5, "height": 5, "iters": 1}], panel_chunks=4)
