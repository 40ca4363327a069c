"""Reporting helpers of bench.py, on the CPU.

bench.reference_residuals reports the reference's own convergence measures
(ComputeFp64Residuals, src/gpu/pdhg_solver.cu:107-161) beside the relative
KKT error; it must agree with the oracle's restatement of that function on
the same FP32 points (bound-violating, interior and at-bound x; both signs
of y; a maximisation LP), and cpu_threads respects OMP_NUM_THREADS.
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def oracle_residuals(oracle, lp, x32, y32, tol):
    L = oracle.lib()
    c32 = (np.where(lp.sense == 1, -1.0, 1.0) * lp.c).astype(np.float32)
    v32 = lp.values.astype(np.float32)
    pv, dv = C.c_double(), C.c_double()
    s = oracle._lp_struct(lp)
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    L.or_fp64_residuals(C.byref(s), f(x32), f(y32), f(c32), f(v32), tol, C.byref(pv), C.byref(dv))
    return pv.value, dv.value


@pytest.mark.parametrize("sense", [0, 1])
def test_reference_residuals_match_oracle(capi, oracle, sense):
    sys.path.insert(0, str(ROOT))
    import bench
    rng = np.random.default_rng(7 + sense)
    lp = capi.generate(None, seed=11 + sense, m=120, n=300, k_min=2, k_max=6, eq_frac=0.5)
    lp.sense = sense
    x = rng.uniform(-1.0, 12.0, lp.n)
    x[::7] = lp.var_lower[::7]          # at the lower bound
    x[1::9] = lp.var_upper[1::9]        # at the upper bound
    x32 = x.astype(np.float32)
    y32 = rng.standard_normal(lp.m).astype(np.float32)
    for tol in (1e-6, 1e-2):
        got = bench.reference_residuals(lp, x32.astype(np.float64), y32.astype(np.float64), tol)
        want = oracle_residuals(oracle, lp, x32, y32, tol)
        assert got[0] == pytest.approx(want[0], rel=1e-12, abs=1e-12)
        assert got[1] == pytest.approx(want[1], rel=1e-12, abs=1e-12)
        assert want[0] > 0 and want[1] > 0


def test_cpu_threads(monkeypatch):
    sys.path.insert(0, str(ROOT))
    import bench
    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    assert bench.cpu_threads() == min(3, __import__("os").cpu_count())
    monkeypatch.setenv("OMP_NUM_THREADS", "not-a-number")
    assert bench.cpu_threads() >= 1
