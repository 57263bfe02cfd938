"""CPU-side checks of the native libraries: they load and export every
symbol the public headers declare (no GPU work is issued)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(chm_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("header,lib", [("chm_jacobi2d.h", "libchm_jacobi2d.so"),
                                        ("charming_amd.h", "libcharming_amd.so")])
def test_library_exports_declared_symbols(header, lib):
    path = os.path.join(ROOT, "charming_amd", "lib", lib)
    assert os.path.exists(path), f"{path} not built"
    L = ctypes.CDLL(path)
    syms = declared(header)
    assert syms
    for s in syms:
        assert hasattr(L, s), f"{lib} does not export {s}"


def test_jacobi_abi_version_without_gpu():
    import charming_amd
    assert charming_amd.jacobi_lib().chm_jacobi2d_abi_version() == 1


@pytest.mark.parametrize("prog", ["hello", "refnum", "pingpong", "async_wait"])
def test_example_programs_built(prog):
    assert os.path.exists(os.path.join(ROOT, "charming_amd", "lib", f"libchm_{prog}.so"))
    assert os.path.exists(os.path.join(ROOT, "charming_amd", "bin", prog))


def test_raw_pingpong_built():
    assert os.access(os.path.join(ROOT, "charming_amd", "bin", "raw_pingpong"),
                     os.X_OK)
