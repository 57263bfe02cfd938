"""Failure paths of the runtime on the GPU, each forced from the host so that
every kernel exits normally (charming_amd/csrc/examples/failure.hip): a
full inbox with a send spin limit, an overflowing refnum-deferral store, and
a refnum that never matches (stopped by the host watchdog).  Each run must
end with the documented status and error word, and the runtime must work
again afterwards."""
import ctypes
import os
import re

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "charming_amd", "lib")
_loaded = {}
ENV = ("CHM_N_PES", "CHM_RING_BYTES", "CHM_MAX_MSG_BYTES", "CHM_SPIN_LIMIT",
       "CHM_DEFERRED_BYTES", "CHM_TIMEOUT_S", "CHM_STATS")


def _rt():
    if "rt" not in _loaded:
        rt = ctypes.CDLL(os.path.join(LIB, "libcharming_amd.so"),
                         mode=ctypes.RTLD_GLOBAL)
        rt.chm_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
        rt.chm_select_program.argtypes = [ctypes.c_char_p]
        _loaded["rt"] = rt
        for prog in ("failure", "hello"):
            _loaded[prog] = ctypes.CDLL(os.path.join(LIB, f"libchm_{prog}.so"))
    return _loaded["rt"]


def _run(capfd, prog, args, env):
    rt = _rt()
    assert rt.chm_select_program(f"{prog}.hip".encode()) == 0
    saved = {k: os.environ.get(k) for k in ENV}
    os.environ.update({"CHM_TIMEOUT_S": "40", "CHM_STATS": "1"})
    os.environ.update({k: str(v) for k, v in env.items()})
    argv = [prog.encode()] + [str(a).encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    print(f"running {prog} {args} {env}", flush=True)
    capfd.readouterr()
    try:
        rc = rt.chm_main(len(argv), arr)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    out, err = capfd.readouterr()
    stats = dict(re.findall(r"(\w+)=([\d.]+)", out.split("[STATS]")[-1]))
    return rc, out, err, stats


def _hello_still_works(capfd):
    rc, out, err, st = _run(capfd, "hello", [], {"CHM_N_PES": 4})
    assert rc == 0, err[-2000:]
    assert len(re.findall(r"Hello I'm \d+ of 8! Received \d+", out)) == 8


def test_full_ring_hits_spin_limit(capfd):
    # 64 KiB ring, 16 KiB messages: the fifth waits for the held PE
    rc, out, err, st = _run(capfd, "failure", ["ring", 64, 16384],
                            {"CHM_N_PES": 2, "CHM_RING_BYTES": 65536,
                             "CHM_MAX_MSG_BYTES": 32768,
                             "CHM_SPIN_LIMIT": 2000})
    assert rc == 1
    assert "device error 8 (send spin limit reached" in err, err[-2000:]
    assert st["error"] == "8" and int(st["send_spins"]) >= 2000
    _hello_still_works(capfd)


def test_deferred_store_overflow(capfd):
    # 64 KiB store, 8 KiB messages whose refnum is never taken
    rc, out, err, st = _run(capfd, "failure", ["defer", 32, 8192],
                            {"CHM_N_PES": 2, "CHM_DEFERRED_BYTES": 65536})
    assert rc == 1
    assert "device error 3 (refnum deferral store full" in err, err[-2000:]
    assert st["error"] == "3"
    assert 1 <= int(st["deferred"]) <= 8
    _hello_still_works(capfd)


def test_unmatched_refnum_stopped_by_watchdog(capfd):
    rc, out, err, st = _run(capfd, "failure", ["never", 1, 64],
                            {"CHM_N_PES": 2, "CHM_TIMEOUT_S": 2})
    assert rc == 1
    assert "watchdog stopped the run" in err, err[-2000:]
    assert int(st["deferred"]) == 1
    # ERR_HOST_ABORT is what the PEs record on their way out
    assert st["error"] == "6"
    _hello_still_works(capfd)
