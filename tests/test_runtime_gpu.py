"""The persistent message-driven runtime on the GPU, checked the way the
reference checks its examples (run the program, read what it prints), with
exact checks on delivery order and count, refnum deferral/revival, payload
bytes and termination.  Programs run in this process through the C ABI
(chm_select_program + chm_main): a cooperative launch cannot be made from a
second process while this one holds the device's cooperative resources."""
import ctypes
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "charming_amd", "lib")
_loaded = {}


def runtime():
    if "rt" not in _loaded:
        rt = ctypes.CDLL(os.path.join(LIB, "libcharming_amd.so"), mode=ctypes.RTLD_GLOBAL)
        rt.chm_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
        rt.chm_select_program.argtypes = [ctypes.c_char_p]
        rt.chm_last_error.restype = ctypes.c_char_p
        _loaded["rt"] = rt
    return _loaded["rt"]


def run(capfd, prog, *args, n_pes=None):
    rt = runtime()
    if prog not in _loaded:
        _loaded[prog] = ctypes.CDLL(os.path.join(LIB, f"libchm_{prog}.so"))
    assert rt.chm_select_program(f"{prog}.hip".encode()) == 0
    os.environ["CHM_TIMEOUT_S"] = "40"
    os.environ["CHM_STATS"] = "1"
    if n_pes:
        os.environ["CHM_N_PES"] = str(n_pes)
    else:
        os.environ.pop("CHM_N_PES", None)
    argv = [prog.encode()] + [str(a).encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    print(f"running {prog} {args} n_pes={n_pes}", flush=True)
    capfd.readouterr()
    rc = rt.chm_main(len(argv), arr)
    out, err = capfd.readouterr()
    os.environ.pop("CHM_N_PES", None)
    assert rc == 0, out[-2000:] + err[-2000:]
    stats = dict(re.findall(r"(\w+)=([\d.]+)", out.split("[STATS]")[-1]))
    return out, stats


@pytest.mark.parametrize("n_pes", [None, 4, 1])
def test_hello_chain(capfd, n_pes):
    out, st = run(capfd, "hello", n_pes=n_pes)
    lines = re.findall(r"Hello I'm (\d+) of (\d+)! Received (\d+)", out)
    P = int(st["pes"])
    n = 2 * P
    assert len(lines) == n
    for k, (i, tot, got) in enumerate(lines):
        assert int(i) == k and int(tot) == n and int(got) == k
    assert int(st["delivered"]) == n and int(st["sent"]) == n
    assert int(st["deferred"]) == 0 and int(st["error"]) == 0


@pytest.mark.parametrize("n_pes", [None, 2])
def test_refnum_deferral_and_revival(capfd, n_pes):
    out, st = run(capfd, "refnum", n_pes=n_pes)
    got = re.findall(r"Received message with refnum (\d+)", out)
    assert got == ["1", "0"]
    assert int(st["deferred"]) == 1 and int(st["revived"]) == 1
    assert int(st["error"]) == 0


def test_pingpong_exact_payloads(capfd):
    out, st = run(capfd, "pingpong", 8, 1 << 20, 20, 2, 1)
    sizes = re.findall(r"Size (\d+) took ([\d.]+) us", out)
    assert [int(s) for s, _ in sizes] == [8 << k for k in range(18)]
    m = re.search(r"round trips (\d+), payload errors (\d+)", out)
    assert m and int(m.group(1)) == 18 * 22 and int(m.group(2)) == 0
    assert "peer done: payload errors 0" in out
    assert int(st["deferred"]) == 0 and int(st["error"]) == 0


def test_raw_mailbox_pingpong_floor():
    # the latency floor without the runtime (examples/baseline/baseline.cu):
    # two workgroups pass flags through per-PE mailboxes; a standalone
    # program with an ordinary launch, run as a child process
    exe = os.path.join(ROOT, "charming_amd", "bin", "raw_pingpong")
    p = subprocess.run([exe, "200", "5"], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    got = re.findall(r"PEs 0 and (\d+) of (\d+), 200 round trips: ([\d.]+) us "
                     r"per one-way trip, errors (\d+)", p.stdout)
    assert len(got) == 2, p.stdout
    n = int(got[0][1])
    assert [int(g[0]) for g in got] == [1, n // 2]
    for _, _, us, err in got:
        assert int(err) == 0 and 0.0 < float(us) < 100.0


@pytest.mark.parametrize("n_pes", [None, 3])
def test_async_wait_on_device_ivars(capfd, n_pes):
    R = 4
    out, st = run(capfd, "async_wait", R, n_pes=n_pes)
    P = int(st["pes"])
    eq = re.findall(r"Round (\d+): all (\d+) ivars equal, payload errors (\d+)", out)
    assert [(int(r), int(k), int(e)) for r, k, e in eq] == \
        [(r, P - 1, 0) for r in range(1, R + 1)]
    ge = re.findall(r"Round (\d+): counter reached (\d+)", out)
    assert [(int(r), int(c)) for r, c in ge] == \
        [(r, r * (P - 1)) for r in range(1, R + 1)]
    assert f"Async wait done: {R} rounds, errors 0" in out
    # init + one publish per other chare per round + two fired waits per round
    assert int(st["delivered"]) == 1 + (P - 1) * R + 2 * R
    assert int(st["error"]) == 0


def test_async_wait_single_pe(capfd):
    out, st = run(capfd, "async_wait", 3, n_pes=1)
    assert "Async wait done: nothing to wait for" in out
    assert int(st["error"]) == 0
