"""Option pairs pm_pdhg_solve accepts and rejects, checked on the host
before any device is touched: the reference algorithm stops only on the
reference's own test and starts only from its own point
(pdhg_solver.cu:288-297,356), so asking it for relative-KKT termination or a
warm start is an argument error, as are unknown algorithms and
terminations, the batch engine's LDS kernels (PM_SPMV_STREAM,
PM_SPMV_STAGED) and the PDLP-only layouts (PM_SPMV_SLICED,
PM_SPMV_COLBLOCK, PM_SPMV_MERGE) under the reference algorithm.  Valid pairs get past the checks (on a machine without a GPU
they then fail for want of a device, not of arguments)."""
import numpy as np
import pytest

PM_ERR_INVALID_ARG = 1


def tiny(capi):
    return capi.LP.from_dense([[1.0, 1.0]], [1.0, 2.0], [1.0], [1.0], [0.0, 0.0], [10.0, 10.0])


@pytest.mark.parametrize("kw, why", [
    (dict(algorithm="reference", termination="relative_kkt"), "relative KKT needs PM_ALG_PDLP"),
    (dict(algorithm="reference", x0=np.zeros(2)), "warm starts need PM_ALG_PDLP"),
    (dict(algorithm="reference", y0=np.zeros(1)), "warm starts need PM_ALG_PDLP"),
    (dict(algorithm=7), "unknown algorithm"),
    (dict(algorithm="pdlp", termination=5), "unknown termination"),
    (dict(check_frequency=0), "check_frequency"),
    (dict(algorithm="pdlp", spmv_variant="stream"), "batch-engine kernels"),
    (dict(algorithm="pdlp", spmv_variant="staged"), "batch-engine kernels"),
    (dict(algorithm="reference", spmv_variant="sliced"), "need PM_ALG_PDLP"),
    (dict(algorithm="reference", spmv_variant="colblock"), "need PM_ALG_PDLP"),
    (dict(algorithm="reference", spmv_variant="merge"), "need PM_ALG_PDLP"),
])
def test_rejected_pairs(capi, kw, why):
    with pytest.raises(capi.PramaanError) as e:
        capi.solve(tiny(capi), **kw)
    assert e.value.code == PM_ERR_INVALID_ARG and why in str(e.value)


@pytest.mark.parametrize("kw", [
    dict(),
    dict(algorithm="pdlp"),
    dict(algorithm="pdlp", termination="relative_kkt"),
    dict(algorithm="pdlp", x0=np.ones(2), y0=np.zeros(1)),
    dict(algorithm="pdlp", spmv_variant="sliced"),
    dict(algorithm="pdlp", spmv_variant="colblock"),
    dict(algorithm="pdlp", spmv_variant="merge"),
])
def test_accepted_pairs_pass_the_checks(capi, kw):
    if capi.device_count() > 0:
        pytest.skip("runs the solve on a GPU machine; covered by tests/test_gpu_options.py")
    with pytest.raises(capi.PramaanError) as e:
        capi.solve(tiny(capi), **kw)
    assert e.value.code != PM_ERR_INVALID_ARG, str(e.value)
