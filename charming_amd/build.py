"""Builds every native library of charming_amd in-tree with hipcc for gfx950.

Outputs go to charming_amd/lib/ (git-ignored, but shipped with the tree to
the GPU host).  Run as `python -m charming_amd.build`.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
if not os.path.exists(HIPCC):
    HIPCC = shutil.which("hipcc") or HIPCC

# gfx950 only; no FMA contraction (bit-exact FP64 against the host oracle);
# the packed-FP32 instructions are disabled.
DEVICE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result",
    "-I", INCLUDE, "-I", CSRC,
]

# (library name, sources, extra link libraries from this build)
LIBS = [
    ("libcharming_amd.so", ["runtime/host_runtime.hip"], []),
    ("libchm_jacobi2d.so", ["jacobi2d/plain.hip", "jacobi2d/persistent.hip",
                           "jacobi2d/batch.hip"],
     ["charming_amd"]),
    # the example chare programs as loadable libraries (run through chm_main)
    ("libchm_hello.so", ["examples/hello.hip"], ["charming_amd"]),
    ("libchm_refnum.so", ["examples/refnum.hip"], ["charming_amd"]),
    ("libchm_pingpong.so", ["examples/pingpong.hip"], ["charming_amd"]),
    ("libchm_async_wait.so", ["examples/async_wait.hip"], ["charming_amd"]),
    ("libchm_failure.so", ["examples/failure.hip"], ["charming_amd"]),
    # GPU-level builds (one PE spanning every workgroup): SRC@gpu is SRC
    # compiled with CHARMING_GPU_LEVEL
    ("libchm_hello_gpu.so", ["examples/hello.hip@gpu"], ["charming_amd"]),
    ("libchm_pingpong_gpu.so", ["examples/pingpong.hip@gpu"], ["charming_amd"]),
]

# chare-program executables: the program's TU + the library-provided main
EXES = [
    ("hello", ["examples/hello.hip"]),
    ("refnum", ["examples/refnum.hip"]),
    ("pingpong", ["examples/pingpong.hip"]),
    ("async_wait", ["examples/async_wait.hip"]),
    ("hello_gpu", ["examples/hello.hip@gpu"]),
    ("pingpong_gpu", ["examples/pingpong.hip@gpu"]),
]
MAIN_SRC = "runtime/main.cpp"
# standalone programs (their own main, no runtime): the raw mailbox pingpong
# that gives the runtime's message latency its floor
STANDALONE_EXES = [
    ("raw_pingpong", ["examples/raw_pingpong.hip"]),
]
BINDIR = os.path.join(PKG, "bin")


def _obj(src):
    return os.path.join(OBJDIR, src.replace("/", "__").replace("@", "__") + ".o")


def _split(src):
    """'path@gpu' -> (path, extra flags of the GPU-level build)."""
    path, _, variant = src.partition("@")
    return path, (["-DCHARMING_GPU_LEVEL"] if variant == "gpu" else [])


def _compile(src, verbose):
    rel, extra = _split(src)
    path = os.path.join(CSRC, rel)
    cmd = [HIPCC] + DEVICE_FLAGS + extra + ["-c", path, "-o", _obj(src)]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{out}")
    return src


# Shared libraries bind their own symbols (-Bsymbolic).  Every chare
# program instantiates the scheduler kernel template, and programs with the
# same workgroup size instantiate the same name; without it, a library
# loaded after another one that exported that name with RTLD_GLOBAL (as
# ctypes loads libchm_jacobi2d.so) would bind its kernel handle to the other
# library's and launch the other program's scheduler.
SHARED_LINK_FLAGS = ["-shared", "-fPIC", "-Wl,-Bsymbolic"]


def _link(name, srcs, deps, verbose):
    cmd = [HIPCC, "--offload-arch=gfx950"] + SHARED_LINK_FLAGS + [
        "-o", os.path.join(LIBDIR, name)] + [_obj(s) for s in srcs]
    cmd += ["-L", LIBDIR] + [f"-l{d}" for d in deps]
    cmd += ["-Wl,-rpath,$ORIGIN", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed for {name}:\n{p.stdout}{p.stderr}")


def _link_exe(name, srcs, verbose):
    cmd = [HIPCC, "--offload-arch=gfx950", "-o", os.path.join(BINDIR, name)]
    cmd += [_obj(s) for s in srcs] + [_obj(MAIN_SRC)]
    cmd += ["-L", LIBDIR, "-lcharming_amd", "-Wl,-rpath,$ORIGIN/../lib", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed for {name}:\n{p.stdout}{p.stderr}")


def _link_standalone(name, srcs, verbose):
    cmd = [HIPCC, "--offload-arch=gfx950", "-o", os.path.join(BINDIR, name)]
    cmd += [_obj(s) for s in srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed for {name}:\n{p.stdout}{p.stderr}")


# Per-lane stack the runtime gives the scheduler kernel (chm_config
# .stack_bytes default); chare programs' entry-method chains must fit in it.
STACK_BYTES = 4096
# chare programs (translation units that include <charming.h>)
CHARE_PROGRAMS = ["jacobi2d/persistent.hip", "examples/hello.hip",
                  "examples/refnum.hip", "examples/pingpong.hip",
                  "examples/async_wait.hip", "examples/failure.hip",
                  "examples/hello.hip@gpu", "examples/pingpong.hip@gpu"]


_ASM = {}


def _device_asm(src):
    """The device assembly of a chare program (compiled once per process)."""
    import tempfile
    if src not in _ASM:
        rel, extra = _split(src)
        with tempfile.TemporaryDirectory() as d:
            out = os.path.join(d, "k.s")
            cmd = [HIPCC] + DEVICE_FLAGS + extra + [
                "-S", "--cuda-device-only", os.path.join(CSRC, rel), "-o", out]
            p = subprocess.run(cmd, capture_output=True, text=True)
            if p.returncode != 0:
                raise RuntimeError(f"hipcc -S failed on {src}:\n{p.stderr}")
            _ASM[src] = open(out).read()
    return _ASM[src]


def stack_need(src):
    """Bytes of per-lane stack the deepest call chain of a chare program's
    scheduler kernel needs: the kernel's own chain down to the dispatch of
    an entry method, plus the largest entry method (reached through a
    function pointer, so the compiler gives the kernel a dynamic stack and
    cannot size it).  Computed from the frame sizes in the device assembly."""
    import re
    asm = _device_asm(src)
    defs = dict(re.findall(r"\.set (\S+)\.private_seg_size, (.+)", asm))
    memo = {}

    def size(name):
        if name not in memo:
            expr = re.sub(r"([.\w$]+?)\.private_seg_size",
                          lambda m: str(size(m.group(1))), defs.get(name, "0"))
            memo[name] = eval(expr.replace("max(", "max(0,"),
                              {"__builtins__": {}, "max": max})
        return memo[name]
    kernel = max(size(n) for n in defs if "scheduler_kernel" in n)
    entries = [size(n) for n in defs if "entry_method" in n]
    return kernel + max(entries, default=0)


def kernel_registers(src):
    """(threads per PE, vector registers per lane) of a chare program's
    scheduler kernel, from the code object metadata in its device assembly.
    The count is the unified total (architectural + accumulation
    registers): the compiler adds the accumulation registers that functions
    holding indirect calls keep across them to the largest callee's count,
    so a kernel can need more than its workgroup size allows."""
    import re
    asm = _device_asm(src)
    m = re.search(r"\.name:\s+_ZN5charm2rt16scheduler_kernelILi(\d+)EEEvi\n",
                  asm)
    if not m:
        raise RuntimeError(f"no scheduler kernel in {src}")
    v = re.search(r"\.vgpr_count:\s+(\d+)", asm[m.end():m.end() + 4000])
    return int(m.group(1)), int(v.group(1))


def register_budget(threads):
    """Vector registers per lane a workgroup of `threads` may use and still
    be resident on one CU of gfx950 (512 per SIMD lane, four SIMDs, wave64,
    allocated in granules of 8)."""
    waves_per_simd = -(-threads // 256)
    return 512 // waves_per_simd // 8 * 8


def build(verbose=False, jobs=None):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    srcs = sorted({s for _, ss, _ in LIBS for s in ss} |
                  {s for _, ss in EXES for s in ss} |
                  {s for _, ss in STANDALONE_EXES for s in ss} | {MAIN_SRC})
    missing = [s for s in srcs
               if not os.path.exists(os.path.join(CSRC, _split(s)[0]))]
    if missing:
        raise FileNotFoundError(f"missing sources: {missing}")
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for f in cf.as_completed([ex.submit(_compile, s, verbose) for s in srcs]):
            f.result()
    for name, ss, deps in LIBS:  # in order: the runtime first
        _link(name, ss, deps, verbose)
    for name, ss in EXES:
        _link_exe(name, ss, verbose)
    for name, ss in STANDALONE_EXES:
        _link_standalone(name, ss, verbose)
    return LIBDIR


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built into", LIBDIR)
