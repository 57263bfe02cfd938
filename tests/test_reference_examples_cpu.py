"""The reference's own example programs (examples/hello, refnum, pingpong)
compile for gfx950 against the charm:: headers with only the documented
edit of the timing header (oracle/ref_examples.py).  Compile only; skipped
where the reference sources are absent."""
import pytest

from oracle import ref_examples as R

pytestmark = pytest.mark.skipif(not R.available(),
                                reason="reference examples not present")


@pytest.mark.parametrize("name,subs,defs", [
    ("hello", 0, ()), ("refnum", 1, ()), ("pingpong", 15, ()),
    # the reference's default build (no SM_LEVEL): its GPU-level branches
    ("pingpong", 15, ("CHARMING_GPU_LEVEL",))])
def test_reference_example_compiles_unchanged(name, subs, defs):
    rc, log, n_sub = R.compile_example(name, defs)
    assert rc == 0, log[-3000:]
    # the only edits: the chrono include and the cuda::std::chrono spelling
    assert n_sub == subs
    assert "error" not in log.lower()
