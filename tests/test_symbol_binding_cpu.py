"""Every chare program's library keeps its own scheduler kernel.

Each chare program instantiates the runtime's scheduler kernel template, and
programs built for the same workgroup size instantiate the same name.  A
shared library that exported that name and was loaded with RTLD_GLOBAL (as
charming_amd loads libchm_jacobi2d.so) used to capture the kernel handle of
every program library loaded after it: the dynamic linker bound the later
library's reference to the first definition, so chm_run launched the other
program's scheduler, whose constant context pointed at memory freed by that
program's last run, and the GPU faulted.  The libraries are now linked with
-Bsymbolic.  This test loads the libraries in that order in a child process
with the dynamic linker's binding trace on (no GPU work: loading only) and
requires that no library binds a symbol to another program's library."""
import glob
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "charming_amd", "lib")


def _libs():
    progs = sorted(p for p in glob.glob(os.path.join(LIB, "libchm_*.so"))
                   if not p.endswith("libchm_jacobi2d.so"))
    progs += sorted(glob.glob(os.path.join(ROOT, "oracle", "_ref", "lib*.so")))
    return progs


def test_program_libraries_bind_their_own_kernels(tmp_path):
    progs = _libs()
    assert len(progs) >= 7, progs
    code = ("import ctypes, sys\n"
            f"ctypes.CDLL({os.path.join(LIB, 'libcharming_amd.so')!r}, "
            "mode=ctypes.RTLD_GLOBAL)\n"
            f"ctypes.CDLL({os.path.join(LIB, 'libchm_jacobi2d.so')!r}, "
            "mode=ctypes.RTLD_GLOBAL)\n"
            f"for p in {progs!r}:\n"
            "    ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)\n")
    env = dict(os.environ, LD_DEBUG="bindings",
               LD_DEBUG_OUTPUT=str(tmp_path / "ld"))
    subprocess.run([sys.executable, "-c", code], check=True, env=env,
                   timeout=120)
    ours = set(progs) | {os.path.join(LIB, "libchm_jacobi2d.so")}
    bad = []
    for f in glob.glob(str(tmp_path / "ld*")):
        for line in open(f, errors="replace"):
            m = re.search(r"binding file (\S+) \[\d+\] to (\S+) \[\d+\]: "
                          r"normal symbol `([^']+)'", line)
            if not m:
                continue
            src, dst, sym = (os.path.realpath(m.group(1)),
                             os.path.realpath(m.group(2)), m.group(3))
            if src in ours and dst in ours and src != dst:
                bad.append((os.path.basename(src), os.path.basename(dst), sym))
    assert not bad, bad[:5]
