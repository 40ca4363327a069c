"""ctypes binding of the engine's C-ABI (include/pramaan_hip.h,
include/pramaan_gen.h).

The library is loaded from pramaan_amd/lib/libpramaan_hip.so (built in-tree
by pramaan_amd.build).  Loading fails loudly when it is missing: there is no
CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libpramaan_hip.so"

# ---- constants (pramaan_hip.h) ------------------------------------------
PM_OK = 0
PM_ERR_INVALID_ARG, PM_ERR_HIP, PM_ERR_NO_DEVICE, PM_ERR_OUT_OF_MEMORY, PM_ERR_INTERNAL = 1, 2, 3, 4, 5
SOLVE_STATUS = {0: "OPTIMAL", 1: "INFEASIBLE", 2: "UNBOUNDED", 3: "ITERATION_LIMIT",
                4: "NUMERICAL_FAILURE", 5: "WARM_START_REJECTED"}
PM_ALG_REFERENCE, PM_ALG_PDLP = 0, 1
PM_TERM_REFERENCE, PM_TERM_RELATIVE_KKT = 0, 1
SPMV = {"auto": 0, "sequential": 1, "scalar": 2, "vector2": 3, "vector4": 4, "vector8": 5,
        "vector16": 6, "vector32": 7, "vector64": 8, "stream": 9, "sliced": 10,
        "staged": 11, "colblock": 12, "merge": 13}
SPMV_NAMES = {v: k for k, v in SPMV.items()}
PM_KERNEL_SPMV_A, PM_KERNEL_SPMV_AT, PM_KERNEL_ITERATION = 0, 1, 2
INF = 1.0e30

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)
_f = C.POINTER(C.c_float)


class pm_lp(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int32), ("sense", C.c_int32),
                ("offset", C.c_double), ("c", _d), ("row_ptr", _i), ("col_idx", _i), ("values", _d),
                ("row_lower", _d), ("row_upper", _d), ("var_lower", _d), ("var_upper", _d)]


class pm_pdhg_options(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int32), ("check_frequency", C.c_int32),
                ("algorithm", C.c_int32), ("termination", C.c_int32), ("spmv_variant", C.c_int32),
                ("ruiz_iterations", C.c_int32), ("max_refinements", C.c_int32), ("device", C.c_int32),
                ("verbose", C.c_int32), ("primal_weight", C.c_double), ("x0", _d), ("y0", _d),
                ("restart_metric", C.c_int32), ("step_rule", C.c_int32), ("feasibility_polishing", C.c_int32)]


class pm_pdhg_result(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("restarts", C.c_int32),
                ("refinements", C.c_int32), ("objective", C.c_double), ("primal_residual", C.c_double),
                ("dual_residual", C.c_double), ("rel_kkt", C.c_double), ("rel_primal", C.c_double),
                ("rel_dual", C.c_double), ("rel_gap", C.c_double), ("primal_objective", C.c_double),
                ("dual_objective", C.c_double), ("primal_weight", C.c_double), ("device_seconds", C.c_double),
                ("x", _d), ("y", _d), ("row_activity", _d), ("polish_attempts", C.c_int32)]


class pm_gen_params(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("k_min", C.c_int32), ("k_max", C.c_int32),
                ("eq_frac", C.c_double), ("noise", C.c_double), ("support", C.c_double),
                ("var_upper", C.c_double), ("seed", C.c_uint64)]


class pm_rh_params(C.Structure):
    _fields_ = [("periods_total", C.c_int32), ("window", C.c_int32), ("demand_noise", C.c_double),
                ("blend_noise", C.c_double), ("seed", C.c_uint64), ("perturb_seed", C.c_uint64)]


class pm_batch_stats(C.Structure):
    _fields_ = [("device_seconds", C.c_double), ("instance_seconds", _d)]


# Every symbol the headers declare, with (restype, argtypes).
SIGNATURES = {
    "pm_version": (C.c_char_p, []),
    "pm_last_error": (C.c_char_p, []),
    "pm_device_count": (C.c_int, [_i]),
    "pm_pdhg_options_default": (None, [C.POINTER(pm_pdhg_options)]),
    "pm_pdhg_solve": (C.c_int, [C.POINTER(pm_lp), C.POINTER(pm_pdhg_options), C.POINTER(pm_pdhg_result)]),
    "pm_launch_spmv": (C.c_int, [_i, _i, _f, C.c_int32, _f, _f, C.c_int32, C.c_int32, C.c_int32]),
    "pm_ruiz_scale": (C.c_int, [C.c_int32, C.c_int32, _i, _i, _d, _d, _d, _d, _d, _d, C.c_int32, _d, _d]),
    "pm_ruiz_scale_noalloc": (C.c_int, [C.c_int32, C.c_int32, C.c_int32] + [C.c_void_p] * 8 + [C.c_int32]
                              + [C.c_void_p] * 3),
    "pm_launch_spmv_t": (C.c_int, [_i, _i, _f, C.c_int32, _f, _f, C.c_int32, C.c_int32, C.c_int32]),
    "pm_launch_pdhg_iteration": (C.c_int, [_f, _f, _f, _f, _f, C.c_int32, C.c_int32, C.c_float, C.c_float]),
    "pm_ctx_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "pm_ctx_destroy": (C.c_int, [C.c_void_p]),
    "pm_ctx_load": (C.c_int, [C.c_void_p, C.POINTER(pm_lp), C.POINTER(pm_pdhg_options)]),
    "pm_ctx_solve": (C.c_int, [C.c_void_p, C.POINTER(pm_pdhg_result)]),
    "pm_ctx_time_kernel": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _d, _d]),
    "pm_ctx_scaling": (C.c_int, [C.c_void_p, _d, _d]),
    "pm_ctx_spmv_variants": (C.c_int, [C.c_void_p, _i, _i]),
    "pm_ctx_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "pm_batch_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(pm_lp), C.POINTER(pm_pdhg_options),
                                  C.POINTER(C.c_void_p)]),
    "pm_batch_set_warm_start": (C.c_int, [C.c_void_p, C.c_int32, _d, _d]),
    "pm_batch_run": (C.c_int, [C.c_void_p, C.POINTER(pm_pdhg_result), C.POINTER(pm_batch_stats)]),
    "pm_batch_destroy": (C.c_int, [C.c_void_p]),
    "pm_batch_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "pm_batch_learned_warm_start": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, _d]),
    "pm_batch_get_warm_start": (C.c_int, [C.c_void_p, C.c_int32, _d, _d]),
    "pm_batch_warm_start_verdicts": (C.c_int, [C.c_void_p, _i, _d]),
    "pm_batch_set_primal_weight": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "pm_batch_scaling": (C.c_int, [C.c_void_p, C.c_int32, _d, _d]),
    "pm_batch_spmv_variants": (C.c_int, [C.c_void_p, _i, _i]),
    "pm_batch_size_classes": (C.c_int, [C.c_void_p, _i, _i, _i, _i]),
    "pm_batch_set_size_classes": (C.c_int, [C.c_void_p, C.c_int32]),
    "pm_batch_time_kernel": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _d, _d]),
    "pm_batch_solve": (C.c_int, [C.c_int32, C.POINTER(pm_lp), C.POINTER(_d), C.POINTER(_d),
                                 C.POINTER(pm_pdhg_options), C.POINTER(pm_pdhg_result), C.POINTER(pm_batch_stats)]),
    "pm_gen_params_named": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(pm_gen_params)]),
    "pm_gen_synthetic": (C.c_int, [C.POINTER(pm_gen_params), C.POINTER(C.c_void_p)]),
    "pm_gen_view": (C.c_int, [C.c_void_p, C.POINTER(pm_lp), C.POINTER(_d), C.POINTER(_d), _d]),
    "pm_gen_free": (None, [C.c_void_p]),
    "pm_gnn_num_params": (C.c_int32, []),
    "pm_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "pm_host_free": (C.c_int, [C.c_void_p]),
    "pm_gnn_set_segments": (C.c_int, [C.c_void_p, C.c_int32, _i]),
    "pm_gnn_set_omega_targets": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pm_gnn_forward_omega": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pm_gnn_set_omega_head": (C.c_int, [C.c_void_p, C.c_int32]),
    "pm_gnn_create": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pm_gnn_destroy": (C.c_int, [C.c_void_p]),
    "pm_gnn_get_params": (C.c_int, [C.c_void_p, _f]),
    "pm_gnn_set_params": (C.c_int, [C.c_void_p, _f]),
    "pm_gnn_set_graph": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _i, _i, _f, _i, _i, _f, _f, _f]),
    "pm_gnn_forward": (C.c_int, [C.c_void_p, _f, _f]),
    "pm_gnn_lp_features": (C.c_int, [C.POINTER(pm_lp), C.c_int32, C.c_int32, _i, _i, _f, _f, _f, _f, _d]),
    "pm_gnn_gradient": (C.c_int, [C.c_void_p, _f, _f, _f, _f]),
    "pm_gnn_train_step": (C.c_int, [C.c_void_p, _f, _f, C.c_float, _f]),
    "pm_rh_params_default": (C.c_int, [C.POINTER(pm_rh_params)]),
    "pm_rh_create": (C.c_int, [C.POINTER(pm_rh_params), C.POINTER(C.c_void_p)]),
    "pm_rh_window_size": (C.c_int, [C.c_void_p, _i, _i, _i]),
    "pm_rh_window": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "pm_rh_free": (None, [C.c_void_p]),
}

_lib = None


class PramaanError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pramaan error {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Loads the engine library (once).  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m pramaan_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != PM_OK:
        raise PramaanError(rc, lib().pm_last_error().decode())


def device_count() -> int:
    n = C.c_int32(0)
    lib().pm_device_count(C.byref(n))
    return n.value


def _p(a: np.ndarray | None, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


@dataclass
class LP:
    """An LP in ModelIR layout (include/pramaan/ir.hpp:62-119), FP64."""
    c: np.ndarray
    row_ptr: np.ndarray
    col_idx: np.ndarray
    values: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    var_lower: np.ndarray
    var_upper: np.ndarray
    sense: int = 0  # 0 minimize, 1 maximize
    offset: float = 0.0
    x_star: np.ndarray | None = None
    y_star: np.ndarray | None = None
    p_star: float | None = None
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self.c, self.values = f64(self.c), f64(self.values)
        self.row_lower, self.row_upper = f64(self.row_lower), f64(self.row_upper)
        self.var_lower, self.var_upper = f64(self.var_lower), f64(self.var_upper)
        self.row_ptr, self.col_idx = i32(self.row_ptr), i32(self.col_idx)

    @property
    def m(self) -> int:
        return len(self.row_ptr) - 1

    @property
    def n(self) -> int:
        return len(self.c)

    @property
    def nnz(self) -> int:
        return len(self.values)

    def struct(self) -> pm_lp:
        return pm_lp(self.m, self.n, self.nnz, self.sense, self.offset, _p(self.c, C.c_double),
                     _p(self.row_ptr, C.c_int32), _p(self.col_idx, C.c_int32), _p(self.values, C.c_double),
                     _p(self.row_lower, C.c_double), _p(self.row_upper, C.c_double),
                     _p(self.var_lower, C.c_double), _p(self.var_upper, C.c_double))

    @staticmethod
    def from_dense(A, c, row_lower, row_upper, var_lower, var_upper, sense=0, offset=0.0) -> "LP":
        A = np.asarray(A, dtype=np.float64).reshape(len(row_lower), len(c))
        rp, ci, vals = [0], [], []
        for row in A:
            for j, a in enumerate(row):
                if a != 0.0:
                    ci.append(j)
                    vals.append(a)
            rp.append(len(ci))
        return LP(np.asarray(c, float), np.asarray(rp), np.asarray(ci, dtype=np.int32), np.asarray(vals, float),
                  np.asarray(row_lower, float), np.asarray(row_upper, float), np.asarray(var_lower, float),
                  np.asarray(var_upper, float), sense, offset)


def options(**kw) -> pm_pdhg_options:
    o = pm_pdhg_options()
    lib().pm_pdhg_options_default(C.byref(o))
    alg = kw.pop("algorithm", None)
    if alg is not None:
        o.algorithm = {"reference": PM_ALG_REFERENCE, "pdlp": PM_ALG_PDLP}.get(alg, alg)
    term = kw.pop("termination", None)
    if term is not None:
        o.termination = {"reference": PM_TERM_REFERENCE, "relative_kkt": PM_TERM_RELATIVE_KKT}.get(term, term)
    sr = kw.pop("step_rule", None)
    if sr is not None:
        o.step_rule = {"halpern": 0, "adaptive": 1}.get(sr, sr)
    rm = kw.pop("restart_metric", None)
    if rm is not None:
        o.restart_metric = {"kkt": 0, "gap": 1}.get(rm, rm)
    sv = kw.pop("spmv_variant", None)
    if sv is not None:
        o.spmv_variant = SPMV.get(sv, sv)
    for k, v in kw.items():
        if k in ("x0", "y0"):
            continue
        setattr(o, k, v)
    return o


@dataclass
class Result:
    status: str
    iterations: int
    restarts: int
    refinements: int
    objective: float
    primal_residual: float
    dual_residual: float
    rel_kkt: float
    rel_primal: float
    rel_dual: float
    rel_gap: float
    primal_objective: float
    dual_objective: float
    primal_weight: float
    device_seconds: float
    x: np.ndarray
    y: np.ndarray
    row_activity: np.ndarray
    polish_attempts: int = 0


def _result(r: pm_pdhg_result, x, y, act) -> Result:
    return Result(SOLVE_STATUS[r.status], r.iterations, r.restarts, r.refinements, r.objective,
                  r.primal_residual, r.dual_residual, r.rel_kkt, r.rel_primal, r.rel_dual, r.rel_gap,
                  r.primal_objective, r.dual_objective, r.primal_weight, r.device_seconds, x, y, act,
                  r.polish_attempts)


def _opts_with_warm(lp: LP, kw: dict):
    x0 = kw.get("x0")
    y0 = kw.get("y0")
    o = options(**kw)
    keep = []
    if x0 is not None:
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        keep.append(x0)
        o.x0 = _p(x0, C.c_double)
    if y0 is not None:
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        keep.append(y0)
        o.y0 = _p(y0, C.c_double)
    return o, keep


def solve(lp: LP, **kw) -> Result:
    """pm_pdhg_solve: one LP, uploaded, solved and freed."""
    o, keep = _opts_with_warm(lp, kw)
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    act = np.zeros(lp.m)
    r = pm_pdhg_result()
    r.x, r.y, r.row_activity = _p(x, C.c_double), _p(y, C.c_double), _p(act, C.c_double)
    s = lp.struct()
    _check(lib().pm_pdhg_solve(C.byref(s), C.byref(o), C.byref(r)))
    del keep
    return _result(r, x, y, act)


class Context:
    """pm_ctx_*: an LP resident on the device."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().pm_ctx_create(device, C.byref(self._h)))
        self.lp = None

    def load(self, lp: LP, **kw) -> None:
        o, keep = _opts_with_warm(lp, kw)
        s = lp.struct()
        _check(lib().pm_ctx_load(self._h, C.byref(s), C.byref(o)))
        self.lp = lp

    def solve(self, want_vectors: bool = True) -> Result:
        lp = self.lp
        x = np.zeros(lp.n) if want_vectors else None
        y = np.zeros(lp.m) if want_vectors else None
        r = pm_pdhg_result()
        if want_vectors:
            r.x, r.y = _p(x, C.c_double), _p(y, C.c_double)
        _check(lib().pm_ctx_solve(self._h, C.byref(r)))
        return _result(r, x, y, None)

    def set_graph(self, use_graph: bool) -> None:
        """pm_ctx_set_graph: replay full chunks as a captured graph (default)
        or issue the same launches directly."""
        _check(lib().pm_ctx_set_graph(self._h, 1 if use_graph else 0))

    def time_kernel(self, which: int, reps: int) -> tuple[float, float]:
        ms, b = C.c_double(), C.c_double()
        _check(lib().pm_ctx_time_kernel(self._h, which, reps, C.byref(ms), C.byref(b)))
        return ms.value, b.value

    def scaling(self) -> tuple[np.ndarray, np.ndarray]:
        """Row and column scaling factors of the loaded LP."""
        dr, dc = np.zeros(max(self.lp.m, 1)), np.zeros(self.lp.n)
        _check(lib().pm_ctx_scaling(self._h, _p(dr, C.c_double), _p(dc, C.c_double)))
        return dr[:self.lp.m], dc

    def spmv_variants(self) -> tuple[str, str]:
        a, t = C.c_int32(), C.c_int32()
        _check(lib().pm_ctx_spmv_variants(self._h, C.byref(a), C.byref(t)))
        return SPMV_NAMES[a.value], SPMV_NAMES[t.value]

    def close(self) -> None:
        if self._h:
            lib().pm_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """pm_batch_*: many LPs resident on the device, solved by one graph launch."""

    def __init__(self, lps: list[LP], device: int = 0, **kw):
        self.lps = list(lps)
        o = options(**kw)
        arr = (pm_lp * len(self.lps))(*[lp.struct() for lp in self.lps])
        self._h = C.c_void_p()
        _check(lib().pm_batch_create(device, len(self.lps), arr, C.byref(o), C.byref(self._h)))

    def set_warm_start(self, i: int, x0=None, y0=None) -> None:
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        y0 = None if y0 is None else np.ascontiguousarray(y0, np.float64)
        _check(lib().pm_batch_set_warm_start(self._h, i, _p(x0, C.c_double), _p(y0, C.c_double)))

    def run(self, want_vectors: bool = False):
        """Returns (results, device_seconds, per-instance seconds)."""
        k = len(self.lps)
        res = (pm_pdhg_result * k)()
        xs, ys = [], []
        if want_vectors:
            for i, lp in enumerate(self.lps):
                xs.append(np.zeros(lp.n))
                ys.append(np.zeros(lp.m))
                res[i].x, res[i].y = _p(xs[-1], C.c_double), _p(ys[-1], C.c_double)
        inst = np.zeros(k)
        st = pm_batch_stats(0.0, _p(inst, C.c_double))
        _check(lib().pm_batch_run(self._h, res, C.byref(st)))
        out = [_result(res[i], xs[i] if want_vectors else None, ys[i] if want_vectors else None, None)
               for i in range(k)]
        return out, st.device_seconds, inst

    def learned_warm_start(self, net: "Gnn", period_cols: int = 0, period_rows: int = 0) -> float:
        """Network predictions as the warm starts of every LP, built on the
        device; returns the device seconds of the step."""
        secs = C.c_double()
        _check(lib().pm_batch_learned_warm_start(self._h, net._h, period_cols, period_rows, C.byref(secs)))
        return secs.value

    def get_warm_start(self, i: int) -> tuple[np.ndarray, np.ndarray]:
        lp = self.lps[i]
        x, y = np.zeros(lp.n), np.zeros(lp.m)
        _check(lib().pm_batch_get_warm_start(self._h, i, _p(x, C.c_double), _p(y, C.c_double)))
        return x, y

    def warm_start_verdicts(self) -> tuple[np.ndarray, np.ndarray]:
        """(accepted: 1 used / 0 replaced by cold / -1 none, FP64 rel. KKT of the start used)."""
        acc = np.zeros(len(self.lps), np.int32)
        kkt = np.zeros(len(self.lps))
        _check(lib().pm_batch_warm_start_verdicts(self._h, _p(acc, C.c_int32), _p(kkt, C.c_double)))
        return acc, kkt

    def set_primal_weight(self, i: int, omega: float) -> None:
        _check(lib().pm_batch_set_primal_weight(self._h, i, float(omega)))

    def scaling(self, i: int) -> tuple[np.ndarray, np.ndarray]:
        """Row and column scaling factors of LP i."""
        lp = self.lps[i]
        dr, dc = np.zeros(max(lp.m, 1)), np.zeros(lp.n)
        _check(lib().pm_batch_scaling(self._h, i, _p(dr, C.c_double), _p(dc, C.c_double)))
        return dr[:lp.m], dc

    def size_classes(self) -> dict:
        """pm_batch_size_classes: launches per half iteration (1 or one per
        size class), and per class its LP count and kernels."""
        n = C.c_int32()
        lps = (C.c_int32 * 3)()
        va = (C.c_int32 * 3)()
        vat = (C.c_int32 * 3)()
        _check(lib().pm_batch_size_classes(self._h, C.byref(n), lps, va, vat))
        return {"launches": n.value, "lps": list(lps), "variant_a": [SPMV_NAMES.get(v, v) for v in va],
                "variant_at": [SPMV_NAMES.get(v, v) for v in vat]}

    def set_size_classes(self, launches: int) -> None:
        """pm_batch_set_size_classes: 1 launch over every LP, or 3 (one per
        size class)."""
        _check(lib().pm_batch_set_size_classes(self._h, launches))

    def spmv_variants(self) -> tuple[int, int]:
        a, at = C.c_int32(), C.c_int32()
        _check(lib().pm_batch_spmv_variants(self._h, C.byref(a), C.byref(at)))
        return a.value, at.value

    def set_graph(self, use_graph: bool) -> None:
        _check(lib().pm_batch_set_graph(self._h, 1 if use_graph else 0))

    def time_kernel(self, which: int, reps: int = 20) -> tuple[float, float]:
        """(mean ms per launch, algorithmic bytes per launch) of an iterate kernel."""
        ms, by = C.c_double(), C.c_double()
        _check(lib().pm_batch_time_kernel(self._h, which, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def close(self) -> None:
        if self._h:
            lib().pm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Gnn:
    """pm_gnn_*: the warm-start network (forward, backward, Adam on the GPU)."""

    def __init__(self, device: int = 0, seed: int = 1):
        self._h = C.c_void_p()
        _check(lib().pm_gnn_create(device, seed, C.byref(self._h)))
        self.n_params = lib().pm_gnn_num_params()
        self.m = self.n = 0

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        _check(lib().pm_gnn_get_params(self._h, _p(out, C.c_float)))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        want = lib().pm_gnn_num_params()
        if p.ndim != 1 or p.size != want:
            raise ValueError(f"parameter vector of {p.size} values; this network has {want} "
                             f"(weights saved for a different feature layout?)")
        _check(lib().pm_gnn_set_params(self._h, _p(p, C.c_float)))

    def set_graph(self, g: dict) -> None:
        """g: dict with m, n, rp, ci, ar, cp, ri, ac, xr, xc (see warmstart.graph_of)."""
        a = {k: np.ascontiguousarray(g[k], np.int32 if k in ("rp", "ci", "cp", "ri") else np.float32)
             for k in ("rp", "ci", "ar", "cp", "ri", "ac", "xr", "xc")}
        self._graph = a
        self._nseg = 1
        self.m, self.n = int(g["m"]), int(g["n"])
        _check(lib().pm_gnn_set_graph(self._h, self.m, self.n, len(a["ci"]), _p(a["rp"], C.c_int32),
                                      _p(a["ci"], C.c_int32), _p(a["ar"], C.c_float), _p(a["cp"], C.c_int32),
                                      _p(a["ri"], C.c_int32), _p(a["ac"], C.c_float), _p(a["xr"], C.c_float),
                                      _p(a["xc"], C.c_float)))

    def set_segments(self, col_off) -> None:
        """LP k of the current graph owns the columns [col_off[k], col_off[k+1])."""
        off = np.ascontiguousarray(col_off, np.int32)
        self._nseg = len(off) - 1
        _check(lib().pm_gnn_set_segments(self._h, self._nseg, _p(off, C.c_int32)))

    def set_omega_targets(self, t) -> None:
        """Targets of the primal-weight head, one per LP (None removes them)."""
        if t is None:
            _check(lib().pm_gnn_set_omega_targets(self._h, None))
            return
        t = np.ascontiguousarray(t, np.float32)
        if len(t) != getattr(self, "_nseg", 1):
            raise ValueError("one primal-weight target per LP of the graph")
        _check(lib().pm_gnn_set_omega_targets(self._h, _p(t, C.c_float)))

    def forward_omega(self) -> np.ndarray:
        out = np.zeros(getattr(self, "_nseg", 1), np.float32)
        _check(lib().pm_gnn_forward_omega(self._h, _p(out, C.c_float)))
        return out

    def set_omega_head(self, enabled: bool) -> None:
        _check(lib().pm_gnn_set_omega_head(self._h, 1 if enabled else 0))

    def forward(self):
        x = np.zeros(self.n, np.float32)
        y = np.zeros(self.m, np.float32)
        _check(lib().pm_gnn_forward(self._h, _p(x, C.c_float), _p(y, C.c_float)))
        return x, y

    def gradient(self, tx, ty):
        tx, ty = np.ascontiguousarray(tx, np.float32), np.ascontiguousarray(ty, np.float32)
        loss = C.c_float()
        g = np.zeros(self.n_params, np.float32)
        _check(lib().pm_gnn_gradient(self._h, _p(tx, C.c_float), _p(ty, C.c_float), C.byref(loss),
                                     _p(g, C.c_float)))
        return loss.value, g

    def train_step(self, tx, ty, lr: float) -> float:
        tx, ty = np.ascontiguousarray(tx, np.float32), np.ascontiguousarray(ty, np.float32)
        loss = C.c_float()
        _check(lib().pm_gnn_train_step(self._h, _p(tx, C.c_float), _p(ty, C.c_float), C.c_float(lr),
                                       C.byref(loss)))
        return loss.value

    def close(self):
        if self._h:
            lib().pm_gnn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


CXX_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libpramaan_cxx.so"
_cxx = None


def cxx() -> C.CDLL:
    """Loads the C++ interface library (certificates; include/pramaan_cert.h)."""
    global _cxx
    if _cxx is None:
        if not CXX_LIB_PATH.exists():
            raise ImportError(f"{CXX_LIB_PATH} is missing: build it with `python -m pramaan_amd.build`")
        L = C.CDLL(str(CXX_LIB_PATH))
        L.pm_cert_write.restype = C.c_int
        L.pm_cert_write.argtypes = [C.POINTER(pm_lp), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_char_p,
                                    C.c_double, C.c_char_p, C.c_char_p, C.c_int]
        L.pm_cert_verify.restype = C.c_int
        L.pm_cert_verify.argtypes = [C.POINTER(pm_lp), C.c_char_p, C.c_char_p, C.c_int]
        _cxx = L
    return _cxx


class pm_ladder_options(C.Structure):
    _fields_ = [("residual_threshold", C.c_double), ("gpu_max_iterations", C.c_int32),
                ("gpu_check_frequency", C.c_int32), ("gpu_tolerance", C.c_double), ("gpu_refined_rung", C.c_int32),
                ("refined_max_iterations", C.c_int32)]


class pm_ladder_result(C.Structure):
    _fields_ = [("stage", C.c_int32), ("escalated", C.c_int32), ("escalation_reason", C.c_int32),
                ("gpu_status", C.c_int32), ("gpu_iterations", C.c_int32), ("gpu_primal_residual", C.c_double),
                ("gpu_dual_residual", C.c_double), ("gpu_objective", C.c_double), ("refined_rung_run", C.c_int32),
                ("refined_status", C.c_int32), ("refined_iterations", C.c_int32), ("refined_rel_kkt", C.c_double),
                ("polish_unavailable", C.c_int32), ("final_status", C.c_int32), ("final_objective", C.c_double),
                ("final_primal_residual", C.c_double), ("final_dual_residual", C.c_double),
                ("x", C.POINTER(C.c_double))]


LADDER_STAGES = ("GPU_FP32", "CPU_FP64_POLISH", "GPU_FP64_REFINED")
ESCALATION_REASONS = ("none", "high_primal_residual", "high_dual_residual", "gpu_iteration_limit",
                      "gpu_numerical_failure", "gpu_not_optimal", "gpu_disabled")


def ladder_solve(lp: LP, residual_threshold=1e-6, gpu_max_iterations=100000, gpu_check_frequency=100,
                 gpu_tolerance=1e-6, refined_rung=False, refined_max_iterations=200000) -> dict:
    """pm_ladder_solve (include/pramaan_ladder.h); returns a dict with the
    stage, escalation reason, per-rung figures and the final x."""
    L = cxx()
    if not hasattr(L.pm_ladder_solve, "_bound"):
        L.pm_ladder_solve.restype = C.c_int
        L.pm_ladder_solve.argtypes = [C.POINTER(pm_lp), C.POINTER(pm_ladder_options), C.POINTER(pm_ladder_result),
                                      C.c_char_p, C.c_int]
        L.pm_ladder_solve._bound = True
    o = pm_ladder_options(residual_threshold, gpu_max_iterations, gpu_check_frequency, gpu_tolerance,
                          1 if refined_rung else 0, refined_max_iterations)
    x = np.zeros(lp.n)
    r = pm_ladder_result()
    r.x = _p(x, C.c_double)
    err = C.create_string_buffer(1024)
    s = lp.struct()
    if L.pm_ladder_solve(C.byref(s), C.byref(o), C.byref(r), err, len(err)) != 0:
        raise RuntimeError(err.value.decode())
    out = {f: getattr(r, f) for f, _ in pm_ladder_result._fields_ if f != "x"}
    out["stage"] = LADDER_STAGES[r.stage]
    out["escalation_reason"] = ESCALATION_REASONS[r.escalation_reason]
    for k in ("gpu_status", "refined_status", "final_status"):
        out[k] = SOLVE_STATUS[out[k]]
    out["x"] = x
    return out


def cert_write(lp: LP, x, path, y=None, status="OPTIMAL", tolerance=1e-6) -> None:
    """Writes a V1 (y None) or V2 certificate; raises ValueError on error."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    if len(x) != lp.n or (y is not None and len(y) != lp.m):
        raise ValueError("certificate: x or y has the wrong length")
    y = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    err = C.create_string_buffer(1024)
    s = lp.struct()
    rc = cxx().pm_cert_write(C.byref(s), _p(x, C.c_double), _p(y, C.c_double), status.encode(), float(tolerance),
                             str(path).encode(), err, len(err))
    if rc != 0:
        raise ValueError(err.value.decode())


def cert_verify(lp: LP, path) -> tuple[int, list[str]]:
    """(0 valid | 1 invalid | 2 unreadable, failure messages)."""
    msg = C.create_string_buffer(1 << 16)
    s = lp.struct()
    rc = cxx().pm_cert_verify(C.byref(s), str(path).encode(), msg, len(msg))
    return rc, [l for l in msg.value.decode().split("\n") if l]


def _csr32(row_ptr, col_idx, values):
    return (np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col_idx, np.int32),
            np.ascontiguousarray(values, np.float32))


def launch_spmv(row_ptr, col_idx, values, x, m, n, variant="auto") -> np.ndarray:
    """pm_launch_spmv: y = A x in FP32 on the GPU."""
    rp, ci, v = _csr32(row_ptr, col_idx, values)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(m, np.float32)
    _check(lib().pm_launch_spmv(_p(rp, C.c_int32), _p(ci, C.c_int32), _p(v, C.c_float), len(v),
                                _p(x, C.c_float), _p(y, C.c_float), m, n, SPMV.get(variant, variant)))
    return y


def ruiz_scale(lp: LP, passes: int = 10) -> dict:
    """pm_ruiz_scale (the reference's ruiz_scale, scaling.cpp:29-150) on
    copies of the LP's arrays: the scaled values, c and bounds and the
    cumulative row_scale / col_scale."""
    rp = np.ascontiguousarray(lp.row_ptr, np.int32)
    ci = np.ascontiguousarray(lp.col_idx, np.int32)
    v = np.array(lp.values, np.float64)
    c = np.array(lp.c, np.float64)
    rl, ru = np.array(lp.row_lower, np.float64), np.array(lp.row_upper, np.float64)
    xl, xu = np.array(lp.var_lower, np.float64), np.array(lp.var_upper, np.float64)
    rs, cs = np.zeros(max(lp.m, 1)), np.zeros(max(lp.n, 1))
    _check(lib().pm_ruiz_scale(lp.m, lp.n, _p(rp, C.c_int32), _p(ci, C.c_int32), _p(v, C.c_double),
                               _p(c, C.c_double), _p(rl, C.c_double), _p(ru, C.c_double), _p(xl, C.c_double),
                               _p(xu, C.c_double), passes, _p(rs, C.c_double), _p(cs, C.c_double)))
    return dict(values=v, c=c, row_lower=rl, row_upper=ru, var_lower=xl, var_upper=xu,
                row_scale=rs[:lp.m], col_scale=cs[:lp.n])


def ruiz_scale_noalloc(m: int, n: int, nnz: int, ptrs: dict, passes: int = 10, stream: int = 0) -> None:
    """pm_ruiz_scale_noalloc on caller-owned device arrays, given as device
    addresses (ints) under the keys row_ptr, col_idx, values, c, row_lower,
    row_upper, var_lower, var_upper, row_scale, col_scale; scaled in place."""
    keys = ("row_ptr", "col_idx", "values", "c", "row_lower", "row_upper", "var_lower", "var_upper")
    _check(lib().pm_ruiz_scale_noalloc(m, n, nnz, *[C.c_void_p(ptrs[k]) for k in keys], passes,
                                       C.c_void_p(ptrs["row_scale"]), C.c_void_p(ptrs["col_scale"]),
                                       C.c_void_p(stream)))


def launch_spmv_t(row_ptr, col_idx, values, x, m, n, variant="auto") -> np.ndarray:
    """pm_launch_spmv_t: y = A^T x in FP32 on the GPU."""
    rp, ci, v = _csr32(row_ptr, col_idx, values)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(n, np.float32)
    _check(lib().pm_launch_spmv_t(_p(rp, C.c_int32), _p(ci, C.c_int32), _p(v, C.c_float), len(v),
                                  _p(x, C.c_float), _p(y, C.c_float), m, n, SPMV.get(variant, variant)))
    return y


def launch_pdhg_iteration(x, y, gradient, ax_bar, b, tau, sigma):
    """pm_launch_pdhg_iteration; returns updated copies of x and y."""
    x = np.array(x, np.float32)
    y = np.array(y, np.float32)
    g = np.ascontiguousarray(gradient, np.float32)
    a = np.ascontiguousarray(ax_bar, np.float32)
    bb = np.ascontiguousarray(b, np.float32)
    _check(lib().pm_launch_pdhg_iteration(_p(x, C.c_float), _p(y, C.c_float), _p(g, C.c_float),
                                          _p(a, C.c_float), _p(bb, C.c_float), len(x), len(y),
                                          C.c_float(tau), C.c_float(sigma)))
    return x, y


def _lp_from_gen(h) -> LP:
    s = pm_lp()
    xs, ys = _d(), _d()
    ps = C.c_double()
    _check(lib().pm_gen_view(h, C.byref(s), C.byref(xs), C.byref(ys), C.byref(ps)))
    m, n, nnz = s.m, s.n, s.nnz
    arr = lambda ptr, k: np.ctypeslib.as_array(ptr, (k,)).copy() if k else np.zeros(0)
    xstar = arr(xs, n) if xs and np.isfinite(ps.value) else None
    ystar = arr(ys, m) if ys and np.isfinite(ps.value) else None
    return LP(arr(s.c, n), np.ctypeslib.as_array(s.row_ptr, (m + 1,)).copy(),
              np.ctypeslib.as_array(s.col_idx, (nnz,)).copy() if nnz else np.zeros(0, np.int32),
              arr(s.values, nnz), arr(s.row_lower, m), arr(s.row_upper, m), arr(s.var_lower, n),
              arr(s.var_upper, n), 0, 0.0, xstar, ystar, ps.value if np.isfinite(ps.value) else None)


class RollingHorizon:
    """pm_rh_*: the multi-period blending plan and its windows."""

    def __init__(self, **params):
        p = pm_rh_params()
        _check(lib().pm_rh_params_default(C.byref(p)))
        for k, v in params.items():
            setattr(p, k, v)
        self.params = p
        self._h = C.c_void_p()
        _check(lib().pm_rh_create(C.byref(p), C.byref(self._h)))
        r, c, w = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().pm_rh_window_size(self._h, C.byref(r), C.byref(c), C.byref(w)))
        self.rows_per_period, self.cols_per_period, self.windows = r.value, c.value, w.value

    def window(self, start: int) -> LP:
        g = C.c_void_p()
        _check(lib().pm_rh_window(self._h, start, C.byref(g)))
        try:
            return _lp_from_gen(g)
        finally:
            lib().pm_gen_free(g)

    def __del__(self):
        try:
            if self._h:
                lib().pm_rh_free(self._h)
        except Exception:
            pass


def with_dense_rows(lp: LP, k: int) -> LP:
    """lp with k rows over every column appended (sum of x <= 2 sum x* + 1):
    rows of n entries among rows of a few dozen, never binding at the
    optimum, so x*, y* (zero duals on the new rows) and p* stay those of lp."""
    n = lp.n
    rp = np.concatenate([lp.row_ptr, lp.row_ptr[-1] + n * np.arange(1, k + 1)])
    ci = np.concatenate([lp.col_idx, np.tile(np.arange(n), k)])
    v = np.concatenate([lp.values, np.ones(n * k)])
    cap = 2.0 * float(np.sum(lp.x_star)) + 1.0
    y_star = None if lp.y_star is None else np.concatenate([lp.y_star, np.zeros(k)])
    return LP(lp.c, rp, ci, v, np.concatenate([lp.row_lower, np.full(k, -INF)]),
              np.concatenate([lp.row_upper, np.full(k, cap)]), lp.var_lower, lp.var_upper, lp.sense, lp.offset,
              lp.x_star, y_star, lp.p_star)


class HostArena:
    """Numpy arrays in page-locked host memory (pm_host_alloc), one block per
    name, grown when a larger array is asked for; an array stays valid
    until the next request under the same name or close()."""

    def __init__(self):
        self._blocks: dict[str, tuple[int, int]] = {}

    def array(self, name: str, shape, dtype) -> np.ndarray:
        dt = np.dtype(dtype)
        count = int(np.prod(shape, dtype=np.int64))
        nbytes = max(count * dt.itemsize, 1)
        blk = self._blocks.get(name)
        if blk is None or blk[1] < nbytes:
            if blk is not None:
                _check(lib().pm_host_free(C.c_void_p(blk[0])))
            p = C.c_void_p()
            _check(lib().pm_host_alloc(nbytes, C.byref(p)))
            blk = (p.value, nbytes)
            self._blocks[name] = blk
        buf = (C.c_char * blk[1]).from_address(blk[0])
        return np.frombuffer(buf, dtype=dt, count=count).reshape(shape)

    def close(self) -> None:
        for ptr, _ in self._blocks.values():
            lib().pm_host_free(C.c_void_p(ptr))
        self._blocks.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generate(which: int | None = None, seed: int = 1, **params) -> LP:
    """Synthetic LP with known optimum (pramaan_gen.h).  `which` names a
    workload (0: 2k x 5k, 1: 500k x 1M, 2: 50k x 100k, 3: batch instance);
    keyword parameters override the fields of pm_gen_params."""
    p = pm_gen_params()
    if which is not None:
        _check(lib().pm_gen_params_named(which, seed, C.byref(p)))
    else:
        p.support, p.var_upper, p.eq_frac, p.seed = 0.3, 10.0, 1.0, seed
        p.k_min, p.k_max = 4, 16  # nonzeros per column (clamped to m)
    for k, v in params.items():
        setattr(p, k, v)
    h = C.c_void_p()
    _check(lib().pm_gen_synthetic(C.byref(p), C.byref(h)))
    try:
        s = pm_lp()
        xs, ys = _d(), _d()
        ps = C.c_double()
        _check(lib().pm_gen_view(h, C.byref(s), C.byref(xs), C.byref(ys), C.byref(ps)))
        m, n, nnz = s.m, s.n, s.nnz
        arr = lambda ptr, k: np.ctypeslib.as_array(ptr, (k,)).copy() if k else np.zeros(0)
        lp = LP(arr(s.c, n), np.ctypeslib.as_array(s.row_ptr, (m + 1,)).copy(),
                np.ctypeslib.as_array(s.col_idx, (nnz,)).copy() if nnz else np.zeros(0, np.int32),
                arr(s.values, nnz), arr(s.row_lower, m), arr(s.row_upper, m), arr(s.var_lower, n),
                arr(s.var_upper, n), 0, 0.0, arr(xs, n), arr(ys, m), ps.value)
        return lp
    finally:
        lib().pm_gen_free(h)
