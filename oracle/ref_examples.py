"""Builds the reference's own example chare programs against this runtime.

Test infrastructure only: it shows that the reference's examples compile and
run against the charm:: headers with no change beyond the one edit the
integration notes document (the timing header: `#include <cuda/std/chrono>`
becomes `#include <charming_chrono.h>` and `cuda::std::chrono` becomes
`charm::chrono`).  The sources are read where they lie under
/root/reference/examples/<name>/ (reference examples hello.cu, refnum.cu,
pingpong.cu and their headers); the edited text is written to a temporary
directory outside the repository, compiled with hipcc for gfx950 and linked
with the runtime's library-provided main.  Outputs go to oracle/_ref/ only.

    python -m oracle.ref_examples            # build every example
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = "/root/reference/examples"
OUT = os.path.join(HERE, "_ref")
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(ROOT, "charming_amd", "csrc")
LIBDIR = os.path.join(ROOT, "charming_amd", "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
if not os.path.exists(HIPCC):
    HIPCC = shutil.which("hipcc") or HIPCC

EXAMPLES = ("hello", "refnum", "pingpong")

# The documented source edits, applied to the text and nothing else.
EDITS = (("#include <cuda/std/chrono>", "#include <charming_chrono.h>"),
         ("cuda::std::chrono", "charm::chrono"))

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument", "-I", INCLUDE]


def available():
    return all(os.path.isdir(os.path.join(REF, n)) for n in EXAMPLES)


def edited_copy(name, dst_dir):
    """Writes the example's .cu/.h files with EDITS applied into dst_dir;
    returns (path of the .cu file, number of substitutions made)."""
    src_dir = os.path.join(REF, name)
    n_sub = 0
    main_cu = None
    for f in sorted(os.listdir(src_dir)):
        if not f.endswith((".cu", ".h")):
            continue
        with open(os.path.join(src_dir, f)) as fh:
            text = fh.read()
        for old, new in EDITS:
            n_sub += text.count(old)
            text = text.replace(old, new)
        # hipcc compiles HIP from a .hip name; the text is unchanged.  The
        # program registers under this file name: "<name>_reference.hip"
        # keeps it apart from this repository's own <name>.hip
        out_name = f[:-3] + "_reference.hip" if f.endswith(".cu") else f
        with open(os.path.join(dst_dir, out_name), "w") as fh:
            fh.write(text)
        if f == name + ".cu":
            main_cu = os.path.join(dst_dir, out_name)
    return main_cu, n_sub


def compile_example(name, defines=(), link=False, out=None):
    """Compiles (and with link=True links) one reference example: the
    executable `out` (with the runtime's main) and the loadable library
    lib<out>.so beside it, which registers the program under the source path
    ".../chm_ref_<name>_*/<name>_reference.hip" (for chm_select_program).  Returns
    (returncode, compiler output, substitutions made)."""
    with tempfile.TemporaryDirectory(prefix=f"chm_ref_{name}_") as d:
        cu, n_sub = edited_copy(name, d)
        obj = os.path.join(d, name + ".o")
        cmd = [HIPCC] + FLAGS + [f"-D{x}" for x in defines] + ["-c", cu, "-o", obj]
        p = subprocess.run(cmd, capture_output=True, text=True)
        log = p.stdout + p.stderr
        if p.returncode != 0 or not link:
            return p.returncode, log, n_sub
        mobj = os.path.join(d, "main.o")
        p = subprocess.run([HIPCC] + FLAGS + ["-c", os.path.join(
            CSRC, "runtime", "main.cpp"), "-o", mobj], capture_output=True,
            text=True)
        if p.returncode != 0:
            return p.returncode, log + p.stdout + p.stderr, n_sub
        os.makedirs(OUT, exist_ok=True)
        exe = out or os.path.join(OUT, name)
        rpath = ["-Wl,-rpath,$ORIGIN/../../charming_amd/lib"]
        p = subprocess.run([HIPCC, "--offload-arch=gfx950", "-o", exe, obj,
                            mobj, "-L", LIBDIR, "-lcharming_amd", "-lpthread"]
                           + rpath, capture_output=True, text=True)
        log += p.stdout + p.stderr
        if p.returncode != 0:
            return p.returncode, log, n_sub
        so = os.path.join(os.path.dirname(exe),
                          "lib" + os.path.basename(exe) + ".so")
        # -Bsymbolic: the library's scheduler kernel handle binds to its own
        # definition, not to a same-named one exported by a library loaded
        # before it (charming_amd.build.SHARED_LINK_FLAGS)
        p = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC",
                            "-Wl,-Bsymbolic",
                            "-o", so, obj, "-L", LIBDIR, "-lcharming_amd",
                            "-lpthread"] + rpath, capture_output=True, text=True)
        return p.returncode, log + p.stdout + p.stderr, n_sub


# (output name, example, extra defines): the reference's default build has
# no SM_LEVEL (its GPU-level branches); these headers select the one-PE-per-
# workgroup branches unless CHARMING_GPU_LEVEL is defined
BUILDS = [("hello", "hello", ()), ("refnum", "refnum", ()),
          ("pingpong", "pingpong", ())]


def build_all(verbose=False):
    if not available():
        return []
    built = []
    for out_name, name, defs in BUILDS:
        rc, log, _ = compile_example(name, defs, link=True,
                                     out=os.path.join(OUT, out_name))
        if rc != 0:
            raise RuntimeError(f"reference example {name} failed:\n{log}")
        if verbose:
            print("built", os.path.join(OUT, out_name))
        built.append(out_name)
    return built


if __name__ == "__main__":
    if not available():
        print("reference examples not present; nothing built")
        sys.exit(0)
    build_all(verbose=True)
