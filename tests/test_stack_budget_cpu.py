"""Every chare program's deepest call chain (scheduler kernel down to the
dispatch, plus the largest entry method, which is reached through a function
pointer) fits the per-lane stack the runtime configures for the scheduler
kernel, with room to spare.  Entry methods are indirect calls, so the kernel
gets a dynamic stack whose size the compiler cannot check; a chain larger
than the configured stack overruns it on the GPU.  And every scheduler
kernel's register count lets one PE workgroup be resident on a CU, which
the cooperative launch needs (chm_init refuses the launch otherwise)."""
import pytest

from charming_amd import build as B


@pytest.mark.parametrize("src", B.CHARE_PROGRAMS)
def test_entry_method_chain_fits_stack(src):
    need = B.stack_need(src)
    assert 0 < need <= B.STACK_BYTES // 2, (src, need)


@pytest.mark.parametrize("src", B.CHARE_PROGRAMS)
def test_scheduler_kernel_fits_one_workgroup_per_cu(src):
    threads, vgprs = B.kernel_registers(src)
    assert vgprs <= B.register_budget(threads), (src, threads, vgprs)


def test_register_budget():
    assert [B.register_budget(t) for t in (256, 512, 768, 1024)] == \
        [512, 256, 168, 128]
