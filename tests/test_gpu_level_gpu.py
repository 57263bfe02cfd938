"""GPU-level mode on the GPU (the reference's build without SM_LEVEL,
scheduler.cu:240-333, get.cu:990-1313): one PE spans every workgroup of the
launch, its leader polls and gates, every workgroup runs main, the entry
methods and invoke between grid barriers.  On one GPU there is one PE, so
the chares of hello and pingpong talk through its own inbox.  Checked the way
the workgroup-level programs are: delivery order and count, byte-exact
seeded payloads, exact refnums, and malloc_user / free_user buffers."""
import ctypes
import os
import re

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "charming_amd", "lib")
_loaded = {}


def run(capfd, prog, *args, wgs=None):
    if "rt" not in _loaded:
        rt = ctypes.CDLL(os.path.join(LIB, "libcharming_amd.so"),
                         mode=ctypes.RTLD_GLOBAL)
        rt.chm_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
        rt.chm_select_program.argtypes = [ctypes.c_char_p]
        _loaded["rt"] = rt
    rt = _loaded["rt"]
    if prog not in _loaded:
        _loaded[prog] = ctypes.CDLL(os.path.join(LIB, f"libchm_{prog}_gpu.so"))
    assert rt.chm_select_program(f"{prog}.hip (gpu-level)".encode()) == 0
    os.environ["CHM_TIMEOUT_S"] = "60"
    os.environ["CHM_STATS"] = "1"
    if wgs:
        os.environ["CHM_GPU_WGS"] = str(wgs)
    argv = [prog.encode()] + [str(a).encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    print(f"running GPU-level {prog} {args} workgroups={wgs}", flush=True)
    capfd.readouterr()
    try:
        rc = rt.chm_main(len(argv), arr)
    finally:
        os.environ.pop("CHM_GPU_WGS", None)
    out, err = capfd.readouterr()
    assert rc == 0, out[-2000:] + err[-2000:]
    stats = dict(re.findall(r"(\w+)=([\d.]+)", out.split("[STATS]")[-1]))
    return out, stats


@pytest.mark.parametrize("wgs", [None, 8, 1])
def test_gpu_level_hello(capfd, wgs):
    out, st = run(capfd, "hello", wgs=wgs)
    assert st["pes"] == "1"
    if wgs:
        assert st["wgs"] == str(wgs)
    lines = re.findall(r"Hello I'm (\d+) of (\d+)! Received (\d+)", out)
    assert [(int(a), int(b), int(c)) for a, b, c in lines] == [(0, 2, 0), (1, 2, 1)]
    assert int(st["delivered"]) == 2 and st["error"] == "0"


@pytest.mark.parametrize("wgs", [None, 16])
def test_gpu_level_pingpong_exact_payloads(capfd, wgs):
    out, st = run(capfd, "pingpong", 8, 1 << 20, 20, 2, 1, wgs=wgs)
    sizes = re.findall(r"Size (\d+) took ([\d.]+) us", out)
    assert [int(s) for s, _ in sizes] == [8 << k for k in range(18)]
    m = re.search(r"round trips (\d+), payload errors (\d+)", out)
    assert m and int(m.group(1)) == 18 * 22 and int(m.group(2)) == 0
    assert "peer done: payload errors 0" in out
    # init (2) + init_done (2) + two messages per round trip
    assert int(st["delivered"]) == 4 + 2 * 18 * 22
    assert int(st["deferred"]) == 0 and st["error"] == "0"
