"""ctypes loader of oracle/liboracle.so.  TEST INFRASTRUCTURE: imported only
by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, as the
checker (or the timed CPU baseline), never by the product path.

The C functions restate the reference (see pdhg_oracle.h for file:line).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_d, _i, _f = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_float)


class or_lp(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int32), ("sense", C.c_int32),
                ("offset", C.c_double), ("c", _d), ("row_ptr", _i), ("col_idx", _i), ("values", _d),
                ("row_lower", _d), ("row_upper", _d), ("var_lower", _d), ("var_upper", _d)]


class or_result(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("restarts", C.c_int32),
                ("objective", C.c_double), ("primal_residual", C.c_double), ("dual_residual", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run python -m pramaan_amd.build")
        L = C.CDLL(str(LIB_PATH))
        L.or_spmv_f32.argtypes = [C.c_int32, _i, _i, _f, _f, _f]
        L.or_spmv_t_f32.argtypes = [C.c_int32, C.c_int32, _i, _i, _f, _f, _f]
        L.or_dual_update.argtypes = [C.c_int32, _f, _f, _f, _f, _f]
        L.or_primal_update.argtypes = [C.c_int32, _f, _f, _f, _f, _f, _f, _f]
        L.or_step1_iteration.argtypes = [_f, _f, _f, _f, _f, C.c_int32, C.c_int32, C.c_float, C.c_float]
        L.or_fp64_residuals.argtypes = [C.POINTER(or_lp), _f, _f, _f, _f, C.c_double, _d, _d]
        L.or_pdhg_solve.argtypes = [C.POINTER(or_lp), C.c_double, C.c_int32, C.c_int32, C.POINTER(or_result),
                                    _f, _f]
        L.or_pdhg_solve.restype = C.c_int
        L.or_transpose.argtypes = [C.c_int32, C.c_int32, _i, _i, _d, _i, _i, _d]
        L.or_ruiz_scale.argtypes = [C.c_int32, C.c_int32, _i, _i, _d, _d, _d, _d, _d, _d, C.c_int32, _d, _d]
        L.or_relative_kkt.argtypes = [C.POINTER(or_lp), _d, _d, _d]
        _lib = L
    return _lib


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def _lp_struct(lp):
    return or_lp(lp.m, lp.n, lp.nnz, lp.sense, lp.offset, _p(lp.c, C.c_double), _p(lp.row_ptr, C.c_int32),
                 _p(lp.col_idx, C.c_int32), _p(lp.values, C.c_double), _p(lp.row_lower, C.c_double),
                 _p(lp.row_upper, C.c_double), _p(lp.var_lower, C.c_double), _p(lp.var_upper, C.c_double))


def spmv(row_ptr, col_idx, values, x, m):
    rp, ci = np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col_idx, np.int32)
    v, x = np.ascontiguousarray(values, np.float32), np.ascontiguousarray(x, np.float32)
    y = np.zeros(m, np.float32)
    lib().or_spmv_f32(m, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v, C.c_float), _p(x, C.c_float),
                      _p(y, C.c_float))
    return y


def spmv_t(row_ptr, col_idx, values, x, m, n):
    rp, ci = np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col_idx, np.int32)
    v, x = np.ascontiguousarray(values, np.float32), np.ascontiguousarray(x, np.float32)
    y = np.zeros(n, np.float32)
    lib().or_spmv_t_f32(m, n, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v, C.c_float), _p(x, C.c_float),
                        _p(y, C.c_float))
    return y


def step1_iteration(x, y, gradient, ax_bar, b, tau, sigma):
    x = np.array(x, np.float32)
    y = np.array(y, np.float32)
    g, a, bb = (np.ascontiguousarray(t, np.float32) for t in (gradient, ax_bar, b))
    lib().or_step1_iteration(_p(x, C.c_float), _p(y, C.c_float), _p(g, C.c_float), _p(a, C.c_float),
                             _p(bb, C.c_float), len(x), len(y), C.c_float(tau), C.c_float(sigma))
    return x, y


def dual_update(ax_bar, sigma, l_r, u_r, y):
    y = np.array(y, np.float32)
    a, s, lo, hi = (np.ascontiguousarray(t, np.float32) for t in (ax_bar, sigma, l_r, u_r))
    lib().or_dual_update(len(y), _p(a, C.c_float), _p(s, C.c_float), _p(lo, C.c_float), _p(hi, C.c_float),
                         _p(y, C.c_float))
    return y


def primal_update(aty, tau, c, l_x, u_x, x):
    x = np.array(x, np.float32)
    xb = np.zeros_like(x)
    a, t, cc, lo, hi = (np.ascontiguousarray(q, np.float32) for q in (aty, tau, c, l_x, u_x))
    lib().or_primal_update(len(x), _p(a, C.c_float), _p(t, C.c_float), _p(cc, C.c_float), _p(lo, C.c_float),
                           _p(hi, C.c_float), _p(x, C.c_float), _p(xb, C.c_float))
    return x, xb


def pdhg_solve(lp, tol=1e-6, max_iterations=100000, check_frequency=100):
    """Reference PDHG (pdhg_solver.cu:169-411) on the CPU.  Returns
    (result dict, x f32, y f32)."""
    s = _lp_struct(lp)
    r = or_result()
    x = np.zeros(lp.n, np.float32)
    y = np.zeros(max(lp.m, 1), np.float32)
    lib().or_pdhg_solve(C.byref(s), tol, max_iterations, check_frequency, C.byref(r), _p(x, C.c_float),
                        _p(y, C.c_float))
    res = dict(status={0: "OPTIMAL", 3: "ITERATION_LIMIT", 4: "NUMERICAL_FAILURE"}[r.status],
               iterations=r.iterations, restarts=r.restarts, objective=r.objective,
               primal_residual=r.primal_residual, dual_residual=r.dual_residual)
    return res, x, y[:lp.m]


def relative_kkt(lp, x, y):
    """FP64 relative KKT (rel, rel_p, rel_d, rel_g, p, d, rp_inf); y in the
    reference sign."""
    s = _lp_struct(lp)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y if len(y) else np.zeros(1), np.float64)
    out = np.zeros(7)
    lib().or_relative_kkt(C.byref(s), _p(x, C.c_double), _p(y, C.c_double), _p(out, C.c_double))
    return dict(rel=out[0], rel_p=out[1], rel_d=out[2], rel_g=out[3], p=out[4], d=out[5], rp_inf=out[6])


def transpose(m, n, row_ptr, col_idx, values):
    rp, ci = np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col_idx, np.int32)
    v = np.ascontiguousarray(values, np.float64)
    nnz = len(v)
    trp, tci, tv = np.zeros(n + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz)
    lib().or_transpose(m, n, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v, C.c_double), _p(trp, C.c_int32),
                       _p(tci, C.c_int32), _p(tv, C.c_double))
    return trp, tci, tv


def ruiz_scale(lp, passes=10):
    """Restated ruiz_scale (scaling.cpp:29-150) on copies; returns the scaled
    arrays and the cumulative factors."""
    v, c = lp.values.copy(), lp.c.copy()
    rl, ru, xl, xu = lp.row_lower.copy(), lp.row_upper.copy(), lp.var_lower.copy(), lp.var_upper.copy()
    rs, cs = np.zeros(max(lp.m, 1)), np.zeros(max(lp.n, 1))
    lib().or_ruiz_scale(lp.m, lp.n, _p(lp.row_ptr, C.c_int32), _p(lp.col_idx, C.c_int32), _p(v, C.c_double),
                        _p(c, C.c_double), _p(rl, C.c_double), _p(ru, C.c_double), _p(xl, C.c_double),
                        _p(xu, C.c_double), passes, _p(rs, C.c_double), _p(cs, C.c_double))
    return dict(values=v, c=c, row_lower=rl, row_upper=ru, var_lower=xl, var_upper=xu,
                row_scale=rs[:lp.m], col_scale=cs[:lp.n])
