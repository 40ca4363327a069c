"""Argument checks of the C-ABI entry points that need no device: null
handles and out-of-range values fail with PM_ERR_INVALID_ARG before any HIP
call, and freeing a null page-locked block does nothing."""
import ctypes as C

PM_OK, PM_ERR_INVALID_ARG = 0, 1


def test_null_handles_are_refused(capi):
    lib = capi.lib()
    assert lib.pm_ctx_set_graph(None, 1) == PM_ERR_INVALID_ARG
    assert lib.pm_batch_set_size_classes(None, 1) == PM_ERR_INVALID_ARG
    assert lib.pm_batch_size_classes(None, None, None, None, None) == PM_ERR_INVALID_ARG
    assert lib.pm_host_alloc(16, None) == PM_ERR_INVALID_ARG
    assert lib.pm_host_free(None) == PM_OK


def test_launch_spmv_checks_the_variant(capi):
    import numpy as np
    import pytest
    rp = np.array([0, 1], np.int32)
    ci = np.array([0], np.int32)
    v = np.array([1.0], np.float32)
    with pytest.raises(capi.PramaanError) as e:
        capi.launch_spmv(rp, ci, v, np.ones(1, np.float32), 1, 1, 14)  # PM_SPMV_NUM_VARIANTS
    assert e.value.code == PM_ERR_INVALID_ARG
