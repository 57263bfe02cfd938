"""charming_amd: MI355X-native message-driven runtime (Charm++-style chares on a
persistent gfx950 kernel) and its FP64 Jacobi2D hot path.

This module is the host-side Python view of the C ABI declared in
include/chm_jacobi2d.h and include/charming_amd.h.  It loads the in-tree
shared libraries built by charming_amd.build and raises if they are missing:
there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")

FORM_PLAIN, FORM_GRAPH, FORM_PERSISTENT = 0, 1, 2
INIT_SEEDED, INIT_REFERENCE = 0, 1

STATUS = {0: "OK", -1: "INVALID", -2: "HIP", -3: "NOMEM", -4: "UNSUPPORTED",
          -5: "DEVICE", -6: "TIMEOUT", -7: "STATE", -8: "NOGPU"}


class ChmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -m charming_amd.build`")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


class Problem(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("chares_r", C.c_int),
                ("chares_c", C.c_int), ("n_chares", C.c_int), ("iters", C.c_int),
                ("seed", C.c_uint64), ("init_mode", C.c_int),
                ("boundary", C.c_double)]


class Result(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("residual_lo", C.c_uint64),
                ("residual_hi", C.c_uint64), ("hash", C.c_uint64),
                ("iters_done", C.c_int), ("launches", C.c_int),
                ("messages", C.c_longlong), ("deferred", C.c_longlong),
                ("sched_overhead_ms", C.c_double)]

    @property
    def residual(self):
        return (self.residual_hi << 64) | self.residual_lo


class Tuning(C.Structure):
    _fields_ = [("panel_chunks", C.c_int), ("rows_in_flight", C.c_int),
                ("band_rows", C.c_int), ("wg_waves", C.c_int),
                ("overlap", C.c_int)]


_jac = None


def jacobi_lib():
    global _jac
    if _jac is None:
        L = _load("libchm_jacobi2d.so")
        vp = C.c_void_p
        L.chm_jacobi2d_create.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(vp)]
        L.chm_jacobi2d_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.chm_jacobi2d_reset.argtypes = [vp]
        L.chm_jacobi2d_solve.argtypes = [vp, C.POINTER(Result)]
        L.chm_jacobi2d_read_grid.argtypes = [vp, C.c_void_p, C.c_size_t]
        L.chm_jacobi2d_hash_region.argtypes = [vp, C.c_int, C.c_int, C.c_int,
                                               C.c_int, C.POINTER(C.c_uint64)]
        L.chm_jacobi2d_destroy.argtypes = [vp]
        L.chm_jacobi2d_phase_stats.argtypes = [vp, C.c_void_p]
        L.chm_jacobi2d_last_error.restype = C.c_char_p
        L.chm_jacobi2d_abi_version.restype = C.c_int
        _jac = L
    return _jac


def _check(rc):
    if rc != 0:
        raise ChmError(rc, jacobi_lib().chm_jacobi2d_last_error().decode())


class Jacobi2D:
    """One Jacobi2D problem resident in HBM, solved in one of the forms:
    'plain' (a launch per sweep), 'graph' (the same launches as a hipGraph)
    or 'persistent' (message-driven chares in one persistent launch)."""

    FORMS = {"plain": FORM_PLAIN, "graph": FORM_GRAPH,
             "persistent": FORM_PERSISTENT}

    def __init__(self, width, height, iters, form="plain", seed=0,
                 init_mode=INIT_SEEDED, boundary=1.0, chares=None, n_chares=0):
        L = jacobi_lib()
        cr, cc = chares if chares else (0, 0)
        self.prob = Problem(width, height, cr, cc, n_chares, iters, seed,
                            init_mode, boundary)
        self.width, self.height, self.iters = width, height, iters
        self.form = form
        h = C.c_void_p()
        _check(L.chm_jacobi2d_create(C.byref(self.prob), self.FORMS[form],
                                     C.byref(h)))
        self.h = h

    def set_tuning(self, panel_chunks=0, rows_in_flight=0, band_rows=0,
                   wg_waves=0, overlap=0):
        t = Tuning(panel_chunks, rows_in_flight, band_rows, wg_waves, overlap)
        _check(jacobi_lib().chm_jacobi2d_set_tuning(self.h, C.byref(t)))

    def reset(self):
        _check(jacobi_lib().chm_jacobi2d_reset(self.h))

    def solve(self):
        r = Result()
        _check(jacobi_lib().chm_jacobi2d_solve(self.h, C.byref(r)))
        return r

    def grid(self):
        out = np.empty((self.height + 2, self.width + 2), dtype=np.float64)
        _check(jacobi_lib().chm_jacobi2d_read_grid(self.h, out.ctypes.data,
                                                   out.size))
        return out

    def hash_region(self, r0, r1, c0, c1):
        v = C.c_uint64()
        _check(jacobi_lib().chm_jacobi2d_hash_region(self.h, r0, r1, c0, c1,
                                                     C.byref(v)))
        return v.value

    def phase_stats(self):
        """Per-chare ms spent in sweep/send/wait in the last persistent
        solve: {'mean': [...], 'max': [...]}"""
        out = np.zeros(8, dtype=np.float64)
        _check(jacobi_lib().chm_jacobi2d_phase_stats(self.h, out.ctypes.data))
        keys = ("sweep", "send", "wait", "recv")
        return {"mean_ms": dict(zip(keys, out[:4].round(3).tolist())),
                "max_ms": dict(zip(keys, out[4:].round(3).tolist()))}

    def close(self):
        if getattr(self, "h", None):
            jacobi_lib().chm_jacobi2d_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
