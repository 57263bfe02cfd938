"""The persistent message-driven runtime on the GPU: the reference's own
verification method (run the example programs, read what they print),
with exact checks on delivery order and count, refnum deferral/revival,
payload bytes and termination."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "charming_amd", "bin")


def run(prog, *args, n_pes=None, timeout=60):
    env = dict(os.environ, CHM_TIMEOUT_S="40", CHM_STATS="1")
    if n_pes:
        env["CHM_N_PES"] = str(n_pes)
    print(f"running {prog} {args} n_pes={n_pes}", flush=True)
    p = subprocess.run([os.path.join(BIN, prog)] + [str(a) for a in args],
                       capture_output=True, text=True, timeout=timeout, env=env)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    stats = dict(re.findall(r"(\w+)=([\d.]+)", p.stdout.split("[STATS]")[-1]))
    return p.stdout, stats


@pytest.mark.parametrize("n_pes", [None, 4, 1])
def test_hello_chain(n_pes):
    out, st = run("hello", n_pes=n_pes)
    lines = re.findall(r"Hello I'm (\d+) of (\d+)! Received (\d+)", out)
    P = int(st["pes"])
    n = 2 * P
    assert len(lines) == n
    for k, (i, tot, got) in enumerate(lines):
        assert int(i) == k and int(tot) == n and int(got) == k
    assert int(st["delivered"]) == n and int(st["sent"]) == n
    assert int(st["deferred"]) == 0 and int(st["error"]) == 0


@pytest.mark.parametrize("n_pes", [None, 2])
def test_refnum_deferral_and_revival(n_pes):
    out, st = run("refnum", n_pes=n_pes)
    got = re.findall(r"Received message with refnum (\d+)", out)
    assert got == ["1", "0"]
    assert int(st["deferred"]) == 1 and int(st["revived"]) == 1
    assert int(st["error"]) == 0


def test_pingpong_exact_payloads():
    out, st = run("pingpong", 8, 1 << 20, 20, 2, 1)
    sizes = re.findall(r"Size (\d+) took ([\d.]+) us", out)
    assert [int(s) for s, _ in sizes] == [8 << k for k in range(18)]
    m = re.search(r"round trips (\d+), payload errors (\d+)", out)
    assert m and int(m.group(1)) == 18 * 22 and int(m.group(2)) == 0
    assert "peer done: payload errors 0" in out
    assert int(st["deferred"]) == 0 and int(st["error"]) == 0
