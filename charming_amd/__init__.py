"""charming_amd: MI355X-native message-driven runtime (Charm++-style chares on a
persistent gfx950 kernel) and its FP64 Jacobi2D hot path.

This module is the host-side Python view of the C ABI declared in
include/chm_jacobi2d.h and include/charming_amd.h.  It loads the in-tree
shared libraries built by charming_amd.build and raises if they are missing:
there is no CPU fallback.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# CHM_LIBDIR: load the native libraries from another build directory
LIBDIR = os.environ.get("CHM_LIBDIR") or os.path.join(PKG, "lib")

FORM_PLAIN, FORM_GRAPH, FORM_PERSISTENT = 0, 1, 2
INIT_SEEDED, INIT_REFERENCE = 0, 1

STATUS = {0: "OK", -1: "INVALID", -2: "HIP", -3: "NOMEM", -4: "UNSUPPORTED",
          -5: "DEVICE", -6: "TIMEOUT", -7: "STATE", -8: "NOGPU", -9: "BUSY"}
E_BUSY = -9
SERVE_ENGINE, SERVE_PLAIN = 0, 1


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
                ("overlap", C.c_int), ("depth", C.c_int),
                ("halo_copy", C.c_int)]


class BatchProblem(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("iters", C.c_int),
                ("block", C.c_int), ("tol", C.c_double), ("seed", C.c_uint64),
                ("init_mode", C.c_int), ("boundary", C.c_double)]


class BatchConfig(C.Structure):
    _fields_ = [("wg_threads", C.c_int), ("block", C.c_int),
                ("panel_chunks", C.c_int), ("rows_in_flight", C.c_int),
                ("timeout_s", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchResult(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("tasks", C.c_longlong),
                ("steals", C.c_longlong), ("idle_polls", C.c_longlong),
                ("cell_updates", C.c_longlong), ("n_wg", C.c_int),
                ("wg_threads", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("mode", C.c_int), ("budget_bytes", C.c_size_t),
                ("max_active", C.c_int), ("wg_threads", C.c_int),
                ("block", C.c_int), ("panel_chunks", C.c_int),
                ("rows_in_flight", C.c_int)]


class ServerCompletion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("iters", C.c_int),
                ("hash", C.c_uint64), ("residual_lo", C.c_uint64),
                ("residual_hi", C.c_uint64), ("residual0_lo", C.c_uint64),
                ("residual0_hi", C.c_uint64), ("device_ms", C.c_double),
                ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("active", C.c_int), ("bytes_in_use", C.c_size_t),
                ("budget_bytes", C.c_size_t), ("submitted", C.c_longlong),
                ("completed", C.c_longlong)]


_jac = None
# live handles, closed at interpreter exit (before the native libraries'
# own exit handlers run) so no device state outlives the program
_live = weakref.WeakSet()


@atexit.register
def _close_live():
    for obj in list(_live):
        try:
            if isinstance(obj, Server):
                obj.stop(abort=True)
            else:
                obj.close()
        except Exception:
            pass


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
        L.chm_jacobi2d_chare_phases.argtypes = [vp, C.c_void_p, C.c_size_t,
                                                C.POINTER(C.c_int)]
        L.chm_jacobi2d_hash_region.argtypes = [vp, C.c_int, C.c_int, C.c_int,
                                               C.c_int, C.POINTER(C.c_uint64)]
        L.chm_jacobi2d_destroy.argtypes = [vp]
        L.chm_jacobi2d_phase_stats.argtypes = [vp, C.c_void_p]
        BP, BC, BR = C.POINTER(BatchProblem), C.POINTER(BatchConfig), C.POINTER(BatchResult)
        L.chm_jacobi2d_batch_run.argtypes = [BP, C.c_int, BC, BR, vp, vp, vp]
        L.chm_jacobi2d_batch_plain.argtypes = [BP, C.c_int, C.c_int, BR, vp, vp, vp]
        L.chm_jacobi2d_batch_autotune.argtypes = [BP, C.c_int, C.c_int, BC, vp,
                                                  C.POINTER(C.c_double)]
        L.chm_jacobi2d_server_start.argtypes = [C.POINTER(ServerConfig), C.POINTER(vp)]
        L.chm_jacobi2d_server_submit.argtypes = [vp, BP, C.POINTER(C.c_uint64)]
        L.chm_jacobi2d_server_poll.argtypes = [vp, C.POINTER(ServerCompletion), C.c_int]
        L.chm_jacobi2d_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.chm_jacobi2d_server_stop.argtypes = [vp, C.c_int, C.c_double]
        L.chm_device_memory.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
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
        _live.add(self)

    def set_tuning(self, panel_chunks=0, rows_in_flight=0, band_rows=0,
                   wg_waves=0, overlap=0, depth=0, halo_copy=0):
        t = Tuning(panel_chunks, rows_in_flight, band_rows, wg_waves, overlap,
                   depth, halo_copy)
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

    def chare_phases(self):
        """Per-chare ms of the last persistent solve: an (n_chares, 4) array
        of sweep, send, wait, recv (chares in index order, row-major)."""
        L, n = jacobi_lib(), C.c_int(0)
        _check(L.chm_jacobi2d_chare_phases(self.h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 4), dtype=np.float64)
        _check(L.chm_jacobi2d_chare_phases(self.h, out.ctypes.data, out.size,
                                           C.byref(n)))
        return out

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


# ---------------------------------------------------------------------------
# batched and served forms
# ---------------------------------------------------------------------------
def _problems(probs):
    """A ctypes array of BatchProblem from dicts (width, height, iters and
    optionally block, tol, seed, init_mode, boundary)."""
    arr = (BatchProblem * len(probs))()
    for i, p in enumerate(probs):
        arr[i] = BatchProblem(p["width"], p["height"], p["iters"],
                              p.get("block", 0), p.get("tol", 0.0),
                              p.get("seed", 0), p.get("init_mode", INIT_SEEDED),
                              p.get("boundary", 1.0))
    return arr


def _batch_outputs(n):
    return (np.zeros(n, np.uint64), np.zeros(2 * n, np.uint64),
            np.zeros(n, np.int32))


def _pack(res, hashes, residuals, iters):
    res_int = [int(residuals[2 * i]) | (int(residuals[2 * i + 1]) << 64)
               for i in range(len(hashes))]
    return {"result": res, "hashes": [int(h) for h in hashes],
            "residuals": res_int, "iters": iters.tolist()}


def batch_run(probs, wg_threads=0, block=0, panel_chunks=0, rows_in_flight=0,
              timeout_s=0.0):
    """Solve every problem in one persistent engine launch."""
    arr = _problems(probs)
    cfg = BatchConfig(wg_threads, block, panel_chunks, rows_in_flight, timeout_s)
    res = BatchResult()
    h, r, it = _batch_outputs(len(probs))
    _check(jacobi_lib().chm_jacobi2d_batch_run(
        arr, len(probs), C.byref(cfg), C.byref(res), h.ctypes.data,
        r.ctypes.data, it.ctypes.data))
    return _pack(res, h, r, it)


def batch_plain(probs, use_graph=False):
    """The same problems one after another, one launch per sweep (or one
    hipGraph per problem)."""
    arr = _problems(probs)
    res = BatchResult()
    h, r, it = _batch_outputs(len(probs))
    _check(jacobi_lib().chm_jacobi2d_batch_plain(
        arr, len(probs), 1 if use_graph else 0, C.byref(res), h.ctypes.data,
        r.ctypes.data, it.ctypes.data))
    return _pack(res, h, r, it)


def batch_autotune(probs, sample_iters=0):
    """Returns (config dict, per-problem tile sides, cell updates/s)."""
    arr = _problems(probs)
    cfg = BatchConfig()
    blocks = np.zeros(len(probs), np.int32)
    rate = C.c_double()
    _check(jacobi_lib().chm_jacobi2d_batch_autotune(
        arr, len(probs), sample_iters, C.byref(cfg), blocks.ctypes.data,
        C.byref(rate)))
    return cfg.as_dict(), blocks.tolist(), rate.value


class Server:
    """A Jacobi2D solve server: requests are admitted against an HBM budget
    and solved in the persistent engine (mode 'engine', continuous batching)
    or one after another in the plain per-launch form (mode 'plain')."""

    def __init__(self, mode="engine", budget_bytes=16 << 30, max_active=0,
                 wg_threads=0, block=0, panel_chunks=0, rows_in_flight=0):
        self.mode = mode
        cfg = ServerConfig(SERVE_ENGINE if mode == "engine" else SERVE_PLAIN,
                           budget_bytes, max_active, wg_threads, block,
                           panel_chunks, rows_in_flight)
        h = C.c_void_p()
        _check(jacobi_lib().chm_jacobi2d_server_start(C.byref(cfg), C.byref(h)))
        self.h = h
        _live.add(self)
        self._buf = (ServerCompletion * 256)()

    def submit(self, prob):
        """Ticket, or None when admission is refused for now (BUSY)."""
        arr = _problems([prob])
        t = C.c_uint64()
        rc = jacobi_lib().chm_jacobi2d_server_submit(self.h, arr, C.byref(t))
        if rc == E_BUSY:
            return None
        _check(rc)
        return t.value

    def poll(self):
        n = jacobi_lib().chm_jacobi2d_server_poll(self.h, self._buf, 256)
        if n < 0:
            _check(n)
        out = []
        for i in range(n):
            c = self._buf[i]
            out.append({"ticket": c.ticket, "iters": c.iters, "hash": c.hash,
                        "residual": c.residual_lo | (c.residual_hi << 64),
                        "residual0": c.residual0_lo | (c.residual0_hi << 64),
                        "device_ms": c.device_ms, "latency_ms": c.latency_ms})
        return out

    def stats(self):
        s = ServerStats()
        _check(jacobi_lib().chm_jacobi2d_server_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def stop(self, abort=False, timeout_s=120.0):
        if getattr(self, "h", None):
            h, self.h = self.h, None
            _check(jacobi_lib().chm_jacobi2d_server_stop(h, 1 if abort else 0,
                                                         timeout_s))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop(abort=a[0] is not None)


def device_memory():
    """(free, total) HBM bytes of the current device."""
    f, t = C.c_size_t(), C.c_size_t()
    _check(jacobi_lib().chm_device_memory(C.byref(f), C.byref(t)))
    return f.value, t.value
