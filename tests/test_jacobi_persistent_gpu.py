"""Message-driven Jacobi2D in one persistent launch: bit-exact against the
CPU oracle (monolithic and the threaded message-driven chare runner), with
refnum deferral exercised and the reference-init known answers."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cam():
    import charming_amd
    charming_amd.jacobi_lib()
    return charming_amd


def expected_messages(cr, cc, iters, depth):
    # init + end message per chare, and per exchange round (one per `depth`
    # sweeps, the first sent by init) both directions of every edge and, with
    # halos at least two deep, of every corner
    edges = 2 * (cr * (cc - 1) + cc * (cr - 1))
    corners = 4 * (cr - 1) * (cc - 1) if depth >= 2 else 0
    rounds = -(-iters // depth)
    return 2 * cr * cc + rounds * (edges + corners)


def test_c1_config_matches_chare_oracle(cam):
    # 512^2 interior, 4x4 chares of 128^2, 100 iterations: full grid + residual
    W = H = 512
    ref, res, deferred = O.jacobi_chares_run(W, H, 4, 4, 100, 8, seed=0xC4A40001)
    with cam.Jacobi2D(W, H, 100, form="persistent", seed=0xC4A40001,
                      chares=(4, 4)) as j:
        r = j.solve()
        g = j.grid()
    assert np.array_equal(g, ref)
    assert r.residual == res
    assert r.hash == O.grid_hash(ref)
    # init + end messages, and per exchange round 48 edges and 36 corners
    # (halos as deep as the 128^2 chares allow, up to sixteen: one round per
    # sixteen sweeps)
    assert r.messages == expected_messages(4, 4, 100, 16)


@pytest.mark.parametrize("W,H,cr,cc,iters", [
    (300, 200, 3, 5, 17), (129, 257, 2, 2, 9), (64, 64, 1, 1, 5),
    (1000, 90, 1, 7, 6), (777, 555, 6, 4, 11), (40, 40, 8, 8, 7)])
def test_ragged_decompositions(cam, W, H, cr, cc, iters):
    ref, res = O.jacobi_run(W, H, iters, seed=W * 7 + H)
    with cam.Jacobi2D(W, H, iters, form="persistent", seed=W * 7 + H,
                      chares=(cr, cc)) as j:
        r = j.solve()
        assert np.array_equal(j.grid(), ref)
        assert r.residual == res


def test_reference_init_kat(cam):
    ref, res = O.jacobi_run(96, 64, 3, mode=1)
    with cam.Jacobi2D(96, 64, 3, form="persistent", chares=(2, 3),
                      init_mode=cam.INIT_REFERENCE) as j:
        r = j.solve()
        g = j.grid()
    assert np.array_equal(g, ref)
    # third sweep at the corner cell, by hand: after two sweeps the corner
    # is (4 + 10 + 2 + 10 + 2) * 0.2 = 5.6000000000000005 and both its
    # interior neighbours (2 + 4 + 2 + 10 + 0) * 0.2 = 3.6, so
    # (5.6000000000000005 + 10 + 3.6 + 10 + 3.6) * 0.2
    assert g[1, 1] == 6.560000000000001
    ref1, _ = O.jacobi_run(96, 64, 1, mode=1)
    assert ref1[1, 1] == 4.0 and ref1[1, 2] == 2.0


@pytest.mark.parametrize("nc,u,depth", [(1, 4, 0), (2, 2, 0), (4, 1, 1), (1, 1, 0),
                                        (2, 1, 0), (1, 2, 1), (0, 0, 1),
                                        (4, 2, 1), (2, 4, 1), (8, 1, 1),
                                        (4, 4, 1), (8, 2, 1)])
def test_tunings_and_rerun(cam, nc, u, depth):
    W, H, iters = 1100, 700, 8
    ref, res = O.jacobi_run(W, H, iters, seed=3)
    with cam.Jacobi2D(W, H, iters, form="persistent", seed=3,
                      chares=(3, 2)) as j:
        j.set_tuning(nc, u, depth=depth)
        r = j.solve()
        assert np.array_equal(j.grid(), ref) and r.residual == res
        j.reset()
        r2 = j.solve()  # a second launch on the same runtime
        assert r2.hash == r.hash and r2.residual == res


def test_default_decomposition_uses_all_pes(cam):
    W = H = 2048
    iters = 20
    with cam.Jacobi2D(W, H, iters, form="persistent", seed=9) as j:
        r = j.solve()
        h_full = r.hash
    ref, res = O.jacobi_run(W, H, iters, seed=9)
    assert h_full == O.grid_hash(ref)
    assert r.residual == res


@pytest.mark.parametrize("halo_copy", [0, 1])
@pytest.mark.parametrize("depth", list(range(1, 17)))
@pytest.mark.parametrize("W,H,cr,cc,iters", [
    (300, 200, 3, 5, 17), (257, 129, 4, 3, 10), (64, 64, 2, 2, 1),
    (100, 100, 5, 5, 2), (130, 70, 1, 1, 9), (500, 40, 2, 7, 5)])
def test_halo_depths_bit_exact(cam, depth, halo_copy, W, H, cr, cc, iters):
    # depth = sweeps per exchange; halo_copy 0 = zero-copy halo messages
    # (allocated in the neighbours' inboxes, written by the sweep), 1 = staged
    ref, res = O.jacobi_run(W, H, iters, seed=W + H)
    with cam.Jacobi2D(W, H, iters, form="persistent", seed=W + H,
                      chares=(cr, cc)) as j:
        j.set_tuning(depth=depth, halo_copy=halo_copy)
        r = j.solve()
        assert np.array_equal(j.grid(), ref)
        assert r.residual == res and r.hash == O.grid_hash(ref)
        assert r.messages == expected_messages(cr, cc, iters, depth)


@pytest.mark.parametrize("nc,u", [(1, 2), (1, 4), (2, 2)])
def test_fused_sweep_shapes(cam, nc, u):
    W, H, iters = 1000, 600, 12
    ref, res = O.jacobi_run(W, H, iters, seed=21, mode=1)
    with cam.Jacobi2D(W, H, iters, form="persistent", seed=21, chares=(3, 2),
                      init_mode=cam.INIT_REFERENCE) as j:
        j.set_tuning(nc, u, depth=2)
        r = j.solve()
        assert np.array_equal(j.grid(), ref) and r.residual == res


def deep_cases(D):
    return [
        (64, 3 * D, 3, 2, 2 * D + 1),    # D-row chares (the thinnest allowed)
        (8 * D, 8 * D, 8, 8, 2 * D + 1),  # 64 chares of D x D: rounds D, D, 1
        (129, 257, 2, 2, D + 1),          # rounds of D, 1
        (200, 130, 1, 1, 3 * D - 1),      # one chare: no messages
        (512, 512, 4, 4, 100)]


@pytest.mark.parametrize("depth", list(range(3, 17)))
@pytest.mark.parametrize("case", range(5))
def test_deep_halo_edge_cases(cam, depth, case):
    W, H, cr, cc, iters = deep_cases(depth)[case]
    ref, res = O.jacobi_run(W, H, iters, seed=W * 3 + H + depth)
    with cam.Jacobi2D(W, H, iters, form="persistent", seed=W * 3 + H + depth,
                      chares=(cr, cc)) as j:
        j.set_tuning(depth=depth)
        r = j.solve()
        assert np.array_equal(j.grid(), ref)
        assert r.residual == res and r.hash == O.grid_hash(ref)
        assert r.messages == expected_messages(cr, cc, iters, depth)


@pytest.mark.parametrize("depth,iters,band", [(3, 12, 0), (4, 13, 0), (5, 14, 0),
                                              (6, 12, 0), (5, 11, 64), (6, 13, 7),
                                              (4, 9, 300), (7, 15, -1),
                                              (8, 17, 256), (8, 16, 128),
                                              (9, 20, 0), (11, 23, 64),
                                              (12, 25, 0), (13, 29, -1),
                                              (14, 30, 7), (15, 31, 0),
                                              (16, 33, 0), (16, 48, 256)])
def test_deep_fused_sweeps_reference_init(cam, depth, iters, band):
    W, H = 1000, 600
    ref, res = O.jacobi_run(W, H, iters, seed=23, mode=1)
    with cam.Jacobi2D(W, H, iters, form="persistent", seed=23, chares=(3, 2),
                      init_mode=cam.INIT_REFERENCE) as j:
        j.set_tuning(1, 1, band_rows=band, depth=depth)
        r = j.solve()
        assert np.array_equal(j.grid(), ref) and r.residual == res


@pytest.mark.parametrize("depth,band", [(3, 64), (8, 40), (9, 256), (16, 100)])
@pytest.mark.parametrize("W", [1000, 1001])
def test_deep_fused_ring_windows_both_cells(cam, depth, band, W):
    # banded passes whose ring-column windows run their steady steps with
    # selects: the right ring falls on a lane's second cell for odd chare
    # widths and on its first for even ones
    H, iters = 700, 2 * depth + 1
    ref, res = O.jacobi_run(W, H, iters, seed=29)
    with cam.Jacobi2D(W, H, iters, form="persistent", seed=29,
                      chares=(2, 2)) as j:
        j.set_tuning(1, 1, band_rows=band, depth=depth)
        r = j.solve()
        assert np.array_equal(j.grid(), ref) and r.residual == res


def test_unsupported_fused_shape_is_refused(cam):
    with cam.Jacobi2D(256, 256, 4, form="persistent", seed=1, chares=(2, 2)) as j:
        j.set_tuning(4, 1, depth=2)
        with pytest.raises(cam.ChmError):
            j.solve()
        j.set_tuning(2, 1, depth=3)
        with pytest.raises(cam.ChmError):
            j.solve()
        j.set_tuning(1, 2, depth=4)
        with pytest.raises(cam.ChmError):
            j.solve()
        j.set_tuning(depth=17)
        with pytest.raises(cam.ChmError):
            j.solve()


def test_thin_chares_fall_back_to_depth_one(cam):
    # one-row chares: the automatic depth is 1, and depth 2 is refused
    ref, res = O.jacobi_run(64, 8, 6, seed=5)
    with cam.Jacobi2D(64, 8, 6, form="persistent", seed=5, chares=(8, 2)) as j:
        r = j.solve()
        assert np.array_equal(j.grid(), ref) and r.residual == res
        j.reset()
        j.set_tuning(depth=2)
        with pytest.raises(cam.ChmError):
            j.solve()
    # two-row chares: depth 3 is refused
    with cam.Jacobi2D(64, 8, 6, form="persistent", seed=5, chares=(4, 2)) as j:
        j.set_tuning(depth=3)
        with pytest.raises(cam.ChmError):
            j.solve()


def test_phase_times_per_chare(cam):
    # per-chare sweep/send/wait/recv times of the last solve, and their mean
    # and maximum as phase_stats reports them; automatic split included
    with cam.Jacobi2D(300, 200, 20, form="persistent", seed=3,
                      chares=(3, 5)) as j:
        j.solve()
        ph = j.chare_phases()
        st = j.phase_stats()
        assert ph.shape == (15, 4)
        assert (ph >= 0).all() and (ph[:, 0] > 0).all()
        keys = list(st["mean_ms"])
        for k, key in enumerate(keys):
            assert st["mean_ms"][key] == pytest.approx(ph[:, k].mean(), abs=2e-3)
            assert st["max_ms"][key] == pytest.approx(ph[:, k].max(), abs=2e-3)
    with cam.Jacobi2D(256, 256, 4, form="persistent", seed=3) as j:
        j.solve()
        assert j.chare_phases().shape[1] == 4 and len(j.chare_phases()) >= 1
    with cam.Jacobi2D(64, 64, 4, form="plain", seed=3) as j:
        j.solve()
        with pytest.raises(cam.ChmError):
            j.chare_phases()
