"""Pins of the oracle's or_fp64_residuals (restating ComputeFp64Residuals,
src/gpu/pdhg_solver.cu:107-161) by values worked out by hand on tiny LPs.

Each case fixes one clause of the reference function so that a plausible
slip in the restatement fails here:
  * pv: a row above its upper bound, a row below its lower bound, x below
    its lower / above its upper bound, the maximum over all of them
    (pdhg_solver.cu:124-134);
  * dv: an interior column (|rc|), a column at its lower bound (max(0,-rc)),
    at its upper bound (max(0, rc)), a fixed column (no violation at all),
    the tol band that decides at_lb / at_ub (136-149), and the division by
    max(1, ||c32||_inf) both below and above 1 (151-157);
  * A x uses the FP64 model values while A^T y uses the FP32 device values.
Expected numbers are written out in the comments next to each case.  The
device pv / dv are compared with this function bit for bit in the GPU
tests, so these cases pin them too.
"""
import ctypes as C

import numpy as np
import pytest

INF = 1e30


def residuals(oracle, capi, A, c, rl, ru, xl, xu, x, y, tol, sense=0):
    lp = capi.LP.from_dense(A, c, rl, ru, xl, xu, sense=sense)
    L = oracle.lib()
    c32 = ((-1.0 if sense == 1 else 1.0) * lp.c).astype(np.float32)
    v32 = lp.values.astype(np.float32)
    x32 = np.asarray(x, np.float32)
    y32 = np.asarray(y if len(y) else [0.0], np.float32)
    pv, dv = C.c_double(), C.c_double()
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    s = oracle._lp_struct(lp)
    L.or_fp64_residuals(C.byref(s), f(x32), f(y32), f(c32), f(v32), tol, C.byref(pv), C.byref(dv))
    return pv.value, dv.value


def test_row_above_upper_bound(oracle, capi):
    # A = [1 2], 0 <= Ax <= 3, x = (1, 2): Ax = 5, pv = 2.
    # y = 0, c = (0.5, -0.25): both columns interior of [0, 10], so
    # dv = max(|0.5|, |-0.25|) / max(1, 0.5) = 0.5.
    pv, dv = residuals(oracle, capi, [[1, 2]], [0.5, -0.25], [0], [3], [0, 0], [10, 10], [1, 2], [0], 1e-6)
    assert pv == 2.0
    assert dv == 0.5


def test_row_below_lower_bound(oracle, capi):
    # 6 <= x1 + 2 x2 <= 8 at x = (1, 2): Ax = 5, pv = 6 - 5 = 1.
    pv, _ = residuals(oracle, capi, [[1, 2]], [0, 0], [6], [8], [0, 0], [10, 10], [1, 2], [0], 1e-6)
    assert pv == 1.0


def test_variable_bound_violations_and_maximum(oracle, capi):
    # Rows satisfied (Ax = 0.25 + 0 in [-1, 1]); x1 = 0.25 below l1 = 1 by
    # 0.75; x2 = 12.5 above u2 = 10 by 2.5 -> but A x = 0.25 + 12.5 = 12.75
    # violates the row by 11.75, which is the maximum.
    pv, _ = residuals(oracle, capi, [[1, 1]], [0, 0], [-1], [1], [1, 0], [10, 10], [0.25, 12.5], [0], 1e-6)
    assert pv == 11.75
    # Without the row: the larger of the two bound violations, 2.5.
    pv, _ = residuals(oracle, capi, [[1, 1]], [0, 0], [-INF], [INF], [1, 0], [10, 10], [0.25, 12.5], [0], 1e-6)
    assert pv == 2.5
    # Only x1 violates: 0.75.
    pv, _ = residuals(oracle, capi, [[1, 1]], [0, 0], [-INF], [INF], [1, 0], [10, 10], [0.25, 3], [0], 1e-6)
    assert pv == 0.75


def test_interior_reduced_cost_and_c_norm_above_one(oracle, capi):
    # A = [2 1], y = 0.5: A^T y = (1, 0.5); c = (-3, 1): rc = (-2, 1.5).
    # Both interior: viol = (2, 1.5); ||c||_inf = 3 > 1, so dv = 2 / 3.
    pv, dv = residuals(oracle, capi, [[2, 1]], [-3, 1], [-INF], [INF], [0, 0], [10, 10], [1, 1], [0.5], 1e-6)
    assert pv == 0.0
    assert dv == 2.0 / 3.0


def test_c_norm_below_one_is_not_used(oracle, capi):
    # c = (0.25, 0.125), y = 0: rc = c, interior -> dv = 0.25 / max(1, 0.25) = 0.25
    # (dividing by ||c||_inf would give 1).
    _, dv = residuals(oracle, capi, [[1, 1]], [0.25, 0.125], [-INF], [INF], [0, 0], [10, 10], [1, 1], [0], 1e-6)
    assert dv == 0.25


def test_at_lower_bound(oracle, capi):
    # x = (0, 0) at l = 0.  c = (-1.5, 2), y = 0: rc = (-1.5, 2).
    # Column 1 at_lb with rc < 0 -> viol 1.5; column 2 at_lb with rc > 0 -> 0.
    # dv = 1.5 / max(1, 2) = 0.75.
    _, dv = residuals(oracle, capi, [[1, 1]], [-1.5, 2], [-INF], [INF], [0, 0], [10, 10], [0, 0], [0], 1e-6)
    assert dv == 0.75
    # Only the dual-feasible column at its bound: dv = 0.
    _, dv = residuals(oracle, capi, [[1, 1]], [0.5, 2], [-INF], [INF], [0, 0], [10, 10], [0, 0], [0], 1e-6)
    assert dv == 0.0


def test_at_upper_bound(oracle, capi):
    # x = (10, 10) at u = 10.  c = (0.75, -4), y = 0: rc = (0.75, -4).
    # Column 1 at_ub with rc > 0 -> viol 0.75; column 2 at_ub with rc < 0 -> 0.
    # dv = 0.75 / max(1, 4) = 0.1875.
    _, dv = residuals(oracle, capi, [[1, 1]], [0.75, -4], [-INF], [INF], [0, 0], [10, 10], [10, 10], [0], 1e-6)
    assert dv == 0.1875


def test_fixed_variable_has_no_violation(oracle, capi):
    # l = u = 2 for column 1: at_lb and at_ub both hold, so any rc gives 0.
    # Column 2 interior with rc = 0.5.  dv = 0.5 / max(1, 5) = 0.1.
    _, dv = residuals(oracle, capi, [[1, 1]], [-5, 0.5], [-INF], [INF], [2, 0], [2, 10], [2, 1], [0], 1e-6)
    assert dv == 0.1
    # The same column with its bounds apart is interior: viol = |-5| = 5, dv = 1.
    _, dv = residuals(oracle, capi, [[1, 1]], [-5, 0.5], [-INF], [INF], [0, 0], [4, 10], [2, 1], [0], 1e-6)
    assert dv == 1.0


def test_tol_band_decides_at_bound(oracle, capi):
    # x1 = 0.125 above l1 = 0, rc1 = +0.5 (dual feasible only at the lower bound).
    # tol = 0.25: 0.125 <= 0 + 0.25 -> at_lb -> viol 0, dv = 0.
    # tol = 0.0625: interior -> viol = 0.5, dv = 0.5.
    args = ([[1, 1]], [0.5, 0], [-INF], [INF], [0, 0], [10, 10], [0.125, 1], [0])
    assert residuals(oracle, capi, *args, 0.25)[1] == 0.0
    assert residuals(oracle, capi, *args, 0.0625)[1] == 0.5
    # The same on the upper side: x1 = 9.875, u = 10, rc1 = -0.5.
    args = ([[1, 1]], [-0.5, 0], [-INF], [INF], [0, 0], [10, 10], [9.875, 1], [0])
    assert residuals(oracle, capi, *args, 0.25)[1] == 0.0
    assert residuals(oracle, capi, *args, 0.0625)[1] == 0.5


def test_maximisation_negates_cost(oracle, capi):
    # max c^T x with c = (0.5, 0): the minimisation cost is (-0.5, 0).  At
    # x = (0, 1), column 1 sits at its lower bound with rc = -0.5 -> viol 0.5.
    # As a minimisation the same column has rc = +0.5 -> viol 0.
    args = ([[1, 1]], [0.5, 0], [-INF], [INF], [0, 0], [10, 10], [0, 1], [0], 1e-6)
    assert residuals(oracle, capi, *args, sense=1)[1] == 0.5
    assert residuals(oracle, capi, *args, sense=0)[1] == 0.0


def test_fp64_values_for_ax_and_fp32_values_for_aty(oracle, capi):
    # A = [0.1]: Ax uses 0.1 in FP64, A^T y uses float(0.1) = 0.100000001490116...
    # Row 0 <= Ax <= 0.25 at x = 3: Ax = 0.30000000000000004, pv = Ax - 0.25.
    # y = 1, c = 0: rc = float(0.1) exactly, interior -> dv = float(0.1).
    pv, dv = residuals(oracle, capi, [[0.1]], [0], [0], [0.25], [0], [10], [3], [1], 1e-6)
    assert pv == 0.1 * 3.0 - 0.25
    assert dv == float(np.float32(0.1))
    assert dv != 0.1


def test_dual_uses_every_row(oracle, capi):
    # Two rows: A = [[1, 0], [1, 1]], y = (1, -3): A^T y = (1 - 3, -3) = (-2, -3).
    # c = (2, 3.5): rc = (0, 0.5).  Column 1 interior -> 0; column 2 interior -> 0.5.
    # dv = 0.5 / max(1, 3.5) = 1/7.
    _, dv = residuals(oracle, capi, [[1, 0], [1, 1]], [2, 3.5], [-INF, -INF], [INF, INF], [0, 0], [10, 10],
                      [1, 1], [1, -3], 1e-6)
    assert dv == pytest.approx(0.5 / 3.5, rel=0, abs=1e-16)
