"""Plain per-launch and hipGraph forms of the Jacobi2D sweep on the GPU,
bit-exact against the CPU oracle."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cam():
    import charming_amd
    charming_amd.jacobi_lib()
    return charming_amd


@pytest.mark.parametrize("W,H,iters,form", [
    (64, 64, 10, "plain"), (128, 96, 7, "plain"), (130, 67, 9, "plain"),
    (257, 33, 5, "graph"), (1, 1, 3, "plain"), (1000, 300, 12, "graph"),
    (513, 512, 25, "plain"), (3, 700, 4, "graph")])
def test_plain_matches_oracle(cam, W, H, iters, form):
    ref, res = O.jacobi_run(W, H, iters, seed=99)
    with cam.Jacobi2D(W, H, iters, form=form, seed=99) as j:
        r = j.solve()
        g = j.grid()
    assert np.array_equal(g, ref)
    assert r.residual == res
    assert r.hash == O.grid_hash(ref)


def test_reference_init_kat(cam):
    with cam.Jacobi2D(40, 30, 1, init_mode=cam.INIT_REFERENCE) as j:
        j.solve()
        g = j.grid()
    inner = g[1:-1, 1:-1]
    assert inner[0, 0] == 4.0 and inner[0, 1] == 2.0 and inner[1, 1] == 0.0
    assert inner[-1, -1] == 4.0 and g[0, 5] == 10.0
    ref, _ = O.jacobi_run(40, 30, 1, mode=1)
    assert np.array_equal(g, ref)


@pytest.mark.parametrize("nc,u,band,waves", [
    (1, 1, 7, 1), (1, 8, 64, 4), (2, 2, 33, 2), (4, 4, 128, 8), (8, 1, 16, 4),
    (8, 2, 300, 1), (4, 1, 1, 3)])
def test_tunings_agree(cam, nc, u, band, waves):
    W, H, iters = 1100, 203, 6
    ref, res = O.jacobi_run(W, H, iters, seed=5)
    with cam.Jacobi2D(W, H, iters, seed=5) as j:
        j.set_tuning(nc, u, band, waves)
        r = j.solve()
        assert np.array_equal(j.grid(), ref)
        assert r.residual == res


def test_repeat_and_reset(cam):
    W, H = 300, 200
    ref, _ = O.jacobi_run(W, H, 8, seed=1)
    with cam.Jacobi2D(W, H, 4, form="graph", seed=1) as j:
        j.solve()
        j.solve()  # 8 sweeps in total: state continues
        assert np.array_equal(j.grid(), ref)
        j.reset()
        j.solve()
        j.solve()
        assert np.array_equal(j.grid(), ref)


def test_hash_region(cam):
    W, H = 200, 150
    ref, _ = O.jacobi_run(W, H, 3, seed=4)
    with cam.Jacobi2D(W, H, 3, seed=4) as j:
        j.solve()
        assert j.hash_region(10, 20, 5, 150) == O.grid_hash(ref, 10, 20, 5, 150)
