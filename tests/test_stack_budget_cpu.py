"""Every chare program's deepest call chain (scheduler kernel down to the
dispatch, plus the largest entry method, which is reached through a function
pointer) fits the per-lane stack the runtime configures for the scheduler
kernel, with room to spare.  Entry methods are indirect calls, so the kernel
gets a dynamic stack whose size the compiler cannot check; a chain larger
than the configured stack overruns it on the GPU."""
import pytest

from charming_amd import build as B


@pytest.mark.parametrize("src", B.CHARE_PROGRAMS)
def test_entry_method_chain_fits_stack(src):
    need = B.stack_need(src)
    assert 0 < need <= B.STACK_BYTES // 2, (src, need)
