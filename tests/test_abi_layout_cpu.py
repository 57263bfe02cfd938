"""The ctypes structures of the Python binding have the sizes and field
offsets the C compiler gives the structs of include/chm_jacobi2d.h (checked
by compiling a probe against the header with gcc; no GPU needed)."""
import ctypes as C
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def c_layout(structs):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "chm_jacobi2d.h"',
             'int main(void) {']
    for cname, fields in structs:
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append('return 0; }')
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    return got


def test_struct_layouts_match_header():
    import charming_amd as cam
    pairs = [("chm_jacobi_problem", cam.Problem), ("chm_jacobi_result", cam.Result),
             ("chm_jacobi_tuning", cam.Tuning),
             ("chm_batch_problem", cam.BatchProblem),
             ("chm_batch_config", cam.BatchConfig),
             ("chm_batch_result", cam.BatchResult),
             ("chm_server_config", cam.ServerConfig),
             ("chm_server_completion", cam.ServerCompletion),
             ("chm_server_stats", cam.ServerStats)]
    got = c_layout([(c, [f for f, _ in py._fields_]) for c, py in pairs])
    for cname, py in pairs:
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_workloads_are_deterministic_and_in_range():
    from charming_amd import workloads as WL
    a = WL.batch_problems()
    assert a == WL.batch_problems()
    assert len(a) == 512
    for i, p in enumerate(a):
        assert 64 <= p["width"] <= 4096
        assert p["height"] == (p["width"] if i % 2 == 0 else p["width"] // 2)
        assert 50 <= p["iters"] <= 500
    arr, reqs = WL.serve_requests(1000, 100.0)
    assert arr[0] == 0.0 and all(b >= a for a, b in zip(arr, arr[1:]))
    assert 5.0 < arr[-1] < 15.0  # ~1000 / 100 s
    for r in reqs:
        assert 256 <= r["width"] <= 8192 and 256 <= r["height"] <= 8192
        assert r["tol"] == 1e-6 and r["iters"] == 2000
