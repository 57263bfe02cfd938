"""chm_jacobi2d_solve_noalloc: the plain solve on a caller-owned grid whose
second grid and accumulators live in the library's static device buffer
(nothing allocated per call), against chm_jacobi2d_solve in the plain form on
the same seeded u_0: the grid bit for bit, the hash and the exact residual.
Two calls in a row on one GPU, the second smaller and with an even sweep
count (its result stays in the caller's grid) after an odd one (copied back
from the static buffer), then one on a non-default stream; a grid beyond the
static buffer is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pitch(w):
    return (w + 18 + 15) // 16 * 16


@pytest.fixture(scope="module")
def env():
    import torch

    import charming_amd as cam
    L = cam.jacobi_lib()
    L.chm_jacobi2d_solve_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int,
                                             C.c_int, C.c_void_p, C.c_void_p]
    return cam, L, torch


def _reference(cam, w, h, iters, seed):
    with cam.Jacobi2D(w, h, iters, form="plain", seed=seed) as j:
        u0 = j.grid()
        r = j.solve()
        return u0, j.grid(), (r.hash, r.residual_lo, r.residual_hi)


def _noalloc(env, u0, w, h, iters, stream=None):
    cam, L, torch = env
    P = _pitch(w)
    buf = torch.zeros((h + 2) * P, dtype=torch.float64, device="cuda")
    assert buf.data_ptr() % 128 == 0
    buf.view(h + 2, P)[:, 15:17 + w] = torch.from_numpy(u0).cuda()
    out3 = torch.zeros(3, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    s = stream.cuda_stream if stream is not None else None
    print(f"no-alloc solve {w}x{h}, {iters} sweeps, stream={s}", flush=True)
    rc = L.chm_jacobi2d_solve_noalloc(buf.data_ptr(), w, h, iters, s,
                                      out3.data_ptr())
    assert rc == 0, L.chm_jacobi2d_last_error().decode()
    torch.cuda.synchronize()
    grid = buf.view(h + 2, P)[:, 15:17 + w].cpu().numpy()
    words = tuple(int(x) & (2**64 - 1) for x in out3.cpu().tolist())
    return grid, words


@pytest.mark.parametrize("w,h,iters,seed,own_stream", [
    (700, 500, 37, 11, False),   # odd: the result comes back from scratch
    (130, 67, 8, 12, False),     # smaller, even, after the larger call
    (257, 33, 5, 13, True),      # a non-default stream
])
def test_noalloc_equals_plain_solve(env, w, h, iters, seed, own_stream):
    cam, L, torch = env
    u0, want_grid, want = _reference(cam, w, h, iters, seed)
    stream = torch.cuda.Stream() if own_stream else None
    grid, got = _noalloc(env, u0, w, h, iters, stream)
    assert np.array_equal(grid.view(np.uint64), want_grid.view(np.uint64))
    assert got == want


def test_noalloc_twice_in_a_row_smaller_second(env):
    cam, L, torch = env
    cases = [(1000, 300, 12, 21), (64, 64, 9, 22)]
    refs = [_reference(cam, *c) for c in cases]
    outs = [_noalloc(env, refs[k][0], *cases[k][:3]) for k in range(2)]
    for (u0, want_grid, want), (grid, got) in zip(refs, outs):
        assert np.array_equal(grid.view(np.uint64), want_grid.view(np.uint64))
        assert got == want


def test_noalloc_refuses_grid_beyond_scratch(env):
    cam, L, torch = env
    out3 = torch.zeros(3, dtype=torch.int64, device="cuda")
    buf = torch.zeros(1024, dtype=torch.float64, device="cuda")
    rc = L.chm_jacobi2d_solve_noalloc(buf.data_ptr(), 4096, 4096, 1, None,
                                      out3.data_ptr())
    assert rc == -1
    assert "no-alloc solve holds at most" in L.chm_jacobi2d_last_error().decode()
