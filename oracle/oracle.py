"""ctypes view of the CPU oracle (oracle/jacobi_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and the
CPU-baseline leg of the bench, as the checker.  The product path never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64 = C.c_uint64
dptr = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_splitmix64_next.argtypes = [C.POINTER(u64)]
        L.orc_splitmix64_next.restype = u64
        L.orc_cell_u01.argtypes = [u64, C.c_int64, C.c_int64]
        L.orc_cell_u01.restype = C.c_double
        L.orc_init_grid.argtypes = [C.c_int, C.c_int, u64, C.c_int, C.c_double, dptr]
        L.orc_jacobi_step.argtypes = [C.c_int, C.c_int, dptr, dptr]
        L.orc_residual_fixed.argtypes = [C.c_int, C.c_int, dptr, dptr,
                                         C.POINTER(u64), C.POINTER(u64)]
        L.orc_u128_to_double.argtypes = [u64, u64]
        L.orc_u128_to_double.restype = C.c_double
        L.orc_grid_hash_region.argtypes = [C.c_int, C.c_int, dptr, C.c_int,
                                           C.c_int, C.c_int, C.c_int]
        L.orc_grid_hash_region.restype = u64
        L.orc_jacobi_run.argtypes = [C.c_int, C.c_int, C.c_int, u64, C.c_int,
                                     C.c_double, dptr, C.POINTER(u64), C.POINTER(u64)]
        L.orc_jacobi_solve_tol.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double,
                                           u64, C.c_int, C.c_double, dptr,
                                           C.POINTER(u64), C.POINTER(u64),
                                           C.POINTER(u64), C.POINTER(u64)]
        L.orc_decompose.argtypes = [C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_block_map.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_alloc_size.argtypes = [C.c_int, u64]
        L.orc_alloc_size.restype = u64
        L.orc_jacobi_chares_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.c_int, u64, C.c_int,
                                            C.c_double, C.c_void_p,
                                            C.POINTER(u64), C.POINTER(u64)]
        L.orc_jacobi_chares_run.restype = C.c_long
        L.orc_jacobi_chares_hashes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.c_int, C.c_int, u64, C.c_int,
                                               C.c_double, C.c_void_p,
                                               C.POINTER(u64), C.POINTER(u64)]
        L.orc_jacobi_chares_hashes.restype = C.c_long
        L.orc_last_run_seconds.restype = C.c_double
        _lib = L
    return _lib


def splitmix64_stream(state, n):
    s = u64(state)
    return [lib().orc_splitmix64_next(C.byref(s)) for _ in range(n)]


def cell_u01(seed, row, col):
    return lib().orc_cell_u01(seed, row, col)


def init_grid(W, H, seed=0, mode=0, boundary=1.0):
    g = np.empty((H + 2, W + 2), dtype=np.float64)
    lib().orc_init_grid(W, H, seed, mode, boundary, g)
    return g


def step(grid):
    H, W = grid.shape[0] - 2, grid.shape[1] - 2
    out = grid.copy()
    lib().orc_jacobi_step(W, H, np.ascontiguousarray(grid), out)
    return out


def residual_fixed(a, b):
    """Exact sum of floor((a-b)^2 * 2^56) over the interior, as a Python int."""
    H, W = a.shape[0] - 2, a.shape[1] - 2
    lo, hi = u64(), u64()
    lib().orc_residual_fixed(W, H, np.ascontiguousarray(a), np.ascontiguousarray(b),
                             C.byref(lo), C.byref(hi))
    return (hi.value << 64) | lo.value


def u128_to_double(s):
    return lib().orc_u128_to_double(s & (2**64 - 1), s >> 64)


def grid_hash(grid, r0=None, r1=None, c0=None, c1=None):
    H, W = grid.shape[0] - 2, grid.shape[1] - 2
    r0 = 1 if r0 is None else r0
    r1 = H + 1 if r1 is None else r1
    c0 = 1 if c0 is None else c0
    c1 = W + 1 if c1 is None else c1
    return lib().orc_grid_hash_region(W, H, np.ascontiguousarray(grid), r0, r1, c0, c1)


def jacobi_run(W, H, iters, seed=0, mode=0, boundary=1.0):
    """Monolithic run; returns (grid u_iters, exact residual of last sweep)."""
    out = np.empty((H + 2, W + 2), dtype=np.float64)
    lo, hi = u64(), u64()
    rc = lib().orc_jacobi_run(W, H, iters, seed, mode, boundary, out,
                              C.byref(lo), C.byref(hi))
    assert rc == 0
    return out, (hi.value << 64) | lo.value


def jacobi_solve_tol(W, H, max_iters, tol, seed=0, mode=0, boundary=1.0):
    """Returns (grid, iterations done, S_last, S_0)."""
    out = np.empty((H + 2, W + 2), dtype=np.float64)
    a, b, c, d = u64(), u64(), u64(), u64()
    n = lib().orc_jacobi_solve_tol(W, H, max_iters, tol, seed, mode, boundary, out,
                                   C.byref(a), C.byref(b), C.byref(c), C.byref(d))
    return out, n, (b.value << 64) | a.value, (d.value << 64) | c.value


def decompose(n_chares, W, H):
    cr, cc = C.c_int(), C.c_int()
    rc = lib().orc_decompose(n_chares, W, H, C.byref(cr), C.byref(cc))
    return None if rc != 0 else (cr.value, cc.value)


def block_map(n_chares, n_pes, counts=None):
    loc = np.zeros(n_chares, dtype=np.int32)
    cnt = None if counts is None else np.ascontiguousarray(counts, dtype=np.int32)
    rc = lib().orc_block_map(n_chares, n_pes,
                             None if cnt is None else cnt.ctypes.data,
                             loc.ctypes.data)
    assert rc == 0
    return loc


def alloc_size(msgtype, payload):
    return lib().orc_alloc_size(msgtype, payload)


def jacobi_chares_run(W, H, cr, cc, iters, n_pes, seed=0, mode=0, boundary=1.0,
                      want_grid=True):
    """Message-driven chare run on host threads.
    Returns (grid or None, exact residual, deferred-message count)."""
    out = np.empty((H + 2, W + 2), dtype=np.float64) if want_grid else None
    lo, hi = u64(), u64()
    d = lib().orc_jacobi_chares_run(W, H, cr, cc, iters, n_pes, seed, mode, boundary,
                                    None if out is None else out.ctypes.data,
                                    C.byref(lo), C.byref(hi))
    assert d >= 0, d
    return out, (hi.value << 64) | lo.value, d


def last_run_seconds():
    return lib().orc_last_run_seconds()


def jacobi_chares_hashes(W, H, cr, cc, iters, n_pes, seed=0, mode=0, boundary=1.0):
    hashes = np.zeros(cr * cc, dtype=np.uint64)
    lo, hi = u64(), u64()
    d = lib().orc_jacobi_chares_hashes(W, H, cr, cc, iters, n_pes, seed, mode,
                                       boundary, hashes.ctypes.data,
                                       C.byref(lo), C.byref(hi))
    assert d >= 0, d
    return hashes, (hi.value << 64) | lo.value
