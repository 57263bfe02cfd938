"""The reference's own example programs, built against this runtime from
their unchanged sources (oracle/ref_examples.py: only the timing header is
edited, as the integration notes document), run on the GPU in this process
through the C ABI.  hello exercises the reference's invoke contract: only
thread 0 passes a buffer (a device `new int`), every other thread passes
null.  The reference's refnum program never calls end(), so it runs until
the host watchdog stops it, which checks that path too.  Skipped where the
reference builds (oracle/_ref/) are absent."""
import ctypes
import os
import re

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "charming_amd", "lib")
REF = os.path.join(ROOT, "oracle", "_ref")
_loaded = {}


def _run(capfd, name, timeout_s, n_pes=None):
    path = os.path.join(REF, f"lib{name}.so")
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (reference sources absent)")
    if "rt" not in _loaded:
        rt = ctypes.CDLL(os.path.join(LIB, "libcharming_amd.so"),
                         mode=ctypes.RTLD_GLOBAL)
        rt.chm_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
        rt.chm_select_program.argtypes = [ctypes.c_char_p]
        _loaded["rt"] = rt
    rt = _loaded["rt"]
    if name not in _loaded:
        _loaded[name] = ctypes.CDLL(path)
    assert rt.chm_select_program(f"{name}_reference.hip".encode()) == 0
    os.environ["CHM_TIMEOUT_S"] = str(timeout_s)
    os.environ["CHM_STATS"] = "1"
    if n_pes:
        os.environ["CHM_N_PES"] = str(n_pes)
    argv = [name.encode()]
    arr = (ctypes.c_char_p * 1)(*argv)
    print(f"running reference {name} n_pes={n_pes}", flush=True)
    capfd.readouterr()
    try:
        rc = rt.chm_main(1, arr)
    finally:
        os.environ.pop("CHM_N_PES", None)
        os.environ["CHM_TIMEOUT_S"] = "40"
    out, err = capfd.readouterr()
    stats = dict(re.findall(r"(\w+)=([\d.]+)", out.split("[STATS]")[-1]))
    return rc, out, err, stats


@pytest.mark.parametrize("n_pes", [None, 3])
def test_reference_hello_unchanged(capfd, n_pes):
    rc, out, err, st = _run(capfd, "hello", 40, n_pes)
    assert rc == 0, out[-2000:] + err[-2000:]
    lines = re.findall(r"Hello I'm (\d+) of (\d+)! Received (\d+)", out)
    n = 2 * int(st["pes"])
    assert [(int(a), int(b), int(c)) for a, b, c in lines] == \
        [(i, n, i) for i in range(n)]
    assert int(st["delivered"]) == n and st["error"] == "0"


def test_reference_refnum_unchanged_until_watchdog(capfd):
    rc, out, err, st = _run(capfd, "refnum", 3)
    # it never calls end(): the watchdog stops it (CHM_E_TIMEOUT -> 1)
    assert rc == 1
    assert "watchdog" in err
    got = re.findall(r"Received message with refnum (\d+)", out)
    assert got == ["1", "0"]   # FIFO: in the order chare 0 sent them
