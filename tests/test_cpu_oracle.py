"""CPU tests: the oracle pinned against known answers and closed forms, the
generators, the MPS reader, and the C-ABI library's exported symbols.
No GPU is touched here."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / "golden"
INF = 1e30


def header_symbols(headers=("pramaan_hip.h", "pramaan_gen.h", "pramaan_gnn.h")):
    syms = set()
    for h in (ROOT / "include" / name for name in headers):
        for name in re.findall(r"^\s*(?:[\w\*]+\s+)+\**(pm_\w+)\s*\(", h.read_text(), re.M):
            syms.add(name)
    return syms


def test_library_exports_every_declared_symbol(capi):
    syms = header_symbols()
    assert len(syms) >= 18
    lib = ctypes.CDLL(str(capi.LIB_PATH))
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(capi.SIGNATURES) >= syms
    # The certificate entry points live in the C++ interface library.
    cert = header_symbols(("pramaan_cert.h", "pramaan_ladder.h"))
    cxx = ctypes.CDLL(str(capi.CXX_LIB_PATH))
    assert cert and not [s for s in sorted(cert) if not hasattr(cxx, s)]
    assert {p.name for p in (ROOT / "include").glob("*.h")} == {
        "pramaan_hip.h", "pramaan_gen.h", "pramaan_gnn.h", "pramaan_cert.h", "pramaan_ladder.h"}


def test_no_device_is_reported_not_faked(capi):
    # Without a GPU, the solver must fail loudly instead of falling back.
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible")
    lp = capi.LP.from_dense([[1.0]], [1.0], [1.0], [INF], [0.0], [INF])
    with pytest.raises(capi.PramaanError) as e:
        capi.solve(lp)
    assert e.value.code == capi.PM_ERR_NO_DEVICE


def test_options_defaults_match_reference(capi):
    o = capi.options()
    # PdhgOptions defaults, include/pramaan/gpu/pdhg_solver.hpp:9-13
    assert (o.tolerance, o.max_iterations, o.check_frequency) == (1e-6, 100000, 100)
    assert o.algorithm == capi.PM_ALG_REFERENCE


# ---- oracle: reference PDHG on the reference's known answers ----------------

def known(capi):
    D = capi.LP.from_dense
    return [
        (D([[1, 0], [0, 2], [3, 2]], [3, 5], [-INF] * 3, [4, 12, 18], [0, 0], [INF, INF], sense=1), 36.0),
        (capi.LP(np.array([2.0, -1.0]), np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), np.zeros(0),
                 np.zeros(0), np.array([0.0, -5.0]), np.array([10.0, 5.0]), 0, 10.0), 5.0),
        (D([[1.0]], [1.0], [3.0], [INF], [0.0], [INF]), 3.0),
        (D([[1, 1]], [1, 1], [5.0], [5.0], [0, 0], [INF, INF]), 5.0),
        (D([[1, 1], [1, 0], [0, 1]], [-1, -2], [-INF] * 3, [4, 3, 3], [0, 0], [INF, INF]), -7.0),
        (D([[1, 1]], [0, 0], [2.0], [INF], [0, 0], [INF, INF]), 0.0),
        (D([[1, 1]], [1, 1], [-INF], [10.0], [2, 3], [5, 7]), 5.0),
    ]


def test_oracle_reference_pdhg_known_answers(capi, oracle):
    """tests/test_gpu_pdhg.cpp:263-404 expected optima: 36, 5, 3, 5, -7, 0, 5."""
    for lp, want in known(capi):
        r, x, _ = oracle.pdhg_solve(lp)
        assert r["status"] == "OPTIMAL"
        assert abs(r["objective"] - want) <= 1e-3 + 1e-4 * abs(want)


def test_oracle_reference_pdhg_afiro(capi, oracle):
    """afiro optimum -464.75314286 (tests/test_sparse_revised_simplex.cpp:80)
    with the reference test's options {200000, 500, 1e-5}."""
    from pramaan_amd.mps import read_mps
    lp = read_mps(GOLDEN / "afiro.mps")
    r, x, _ = oracle.pdhg_solve(lp, tol=1e-5, max_iterations=200000, check_frequency=500)
    assert r["status"] == "OPTIMAL"
    assert abs(r["objective"] + 464.75314286) <= 1e-3


def test_oracle_iteration_limit_and_nan(capi, oracle):
    lp = capi.generate(0, seed=1)
    r, _, _ = oracle.pdhg_solve(lp, tol=1e-12, max_iterations=250, check_frequency=100)
    assert r["status"] == "ITERATION_LIMIT" and r["iterations"] == 250
    bad = capi.LP.from_dense([[1.0]], [np.nan], [0.0], [1.0], [0.0], [1.0])
    assert oracle.pdhg_solve(bad)[0]["status"] == "NUMERICAL_FAILURE"


def test_oracle_spmv_against_scipy(capi, oracle):
    lp = capi.generate(0, seed=5)
    A = sp.csr_matrix((lp.values.astype(np.float32), lp.col_idx, lp.row_ptr), shape=(lp.m, lp.n))
    x = np.random.default_rng(0).standard_normal(lp.n).astype(np.float32)
    y = np.random.default_rng(1).standard_normal(lp.m).astype(np.float32)
    assert np.allclose(oracle.spmv(lp.row_ptr, lp.col_idx, lp.values, x, lp.m), A @ x, rtol=1e-5, atol=1e-5)
    assert np.allclose(oracle.spmv_t(lp.row_ptr, lp.col_idx, lp.values, y, lp.m, lp.n), A.T @ y, rtol=1e-5,
                       atol=1e-5)


def test_oracle_transpose_against_scipy(capi, oracle):
    lp = capi.generate(0, seed=2)
    trp, tci, tv = oracle.transpose(lp.m, lp.n, lp.row_ptr, lp.col_idx, lp.values)
    At = sp.csr_matrix((lp.values, lp.col_idx, lp.row_ptr), shape=(lp.m, lp.n)).T.tocsr()
    At.sort_indices()
    assert np.array_equal(trp, At.indptr) and np.array_equal(tci, At.indices) and np.array_equal(tv, At.data)


def test_oracle_updates_closed_form(oracle):
    # dual: w = y/s + a; y' = s (w - clamp(w, l, u)); sigma <= 0 leaves y.
    y = oracle.dual_update([3.0, -1.0, 0.5], [2.0, 0.0, 1.0], [0.0, 0.0, -1.0], [1.0, 0.0, 1.0], [2.0, 7.0, 0.0])
    assert y.tolist() == [2.0 * (4.0 - 1.0), 7.0, 0.0]
    x, xb = oracle.primal_update([1.0, 0.0], [0.5, 1.0], [1.0, -3.0], [0.0, 0.0], [10.0, 2.0], [2.0, 1.0])
    assert x.tolist() == [1.0, 2.0] and xb.tolist() == [0.0, 3.0]


# ---- oracle: relative KKT pinned by the closed-form optimum ----------------

@pytest.mark.parametrize("which,seed", [(0, 1), (0, 2), (3, 5)])
def test_kkt_zero_at_constructed_optimum(capi, oracle, which, seed):
    lp = capi.generate(which, seed=seed)
    k = oracle.relative_kkt(lp, lp.x_star, -lp.y_star)
    assert k["rel"] <= 1e-13
    assert k["p"] == pytest.approx(lp.p_star, rel=1e-12)


def test_kkt_detects_plausible_mistakes(capi, oracle):
    lp = capi.generate(0, seed=1)
    xs, ys = lp.x_star, -lp.y_star
    assert oracle.relative_kkt(lp, xs, lp.y_star)["rel"] > 1e-3          # wrong dual sign
    x2 = xs.copy(); x2[np.argmax(xs)] += 1e-3
    assert oracle.relative_kkt(lp, x2, ys)["rel_p"] > 1e-8                # primal perturbation
    y2 = ys.copy(); y2[0] += 1e-3
    assert oracle.relative_kkt(lp, xs, y2)["rel_g"] > 1e-8                # dual perturbation
    # maximisation flips the cost: same point is optimal for max(-c)
    lpm = capi.LP(-lp.c, lp.row_ptr, lp.col_idx, lp.values, lp.row_lower, lp.row_upper, lp.var_lower,
                  lp.var_upper, 1)
    assert oracle.relative_kkt(lpm, xs, ys)["rel"] <= 1e-13


def test_kkt_inequality_rows_and_free_vars(capi, oracle):
    """min -x1 - 2x2, x1 + x2 <= 4, x1 <= 3, x2 <= 3, x >= 0: x* = (1, 3),
    textbook duals (-1, 0, -1) -> reference sign (1, 0, 1)."""
    lp = capi.LP.from_dense([[1, 1], [1, 0], [0, 1]], [-1, -2], [-INF] * 3, [4, 3, 3], [0, 0], [INF, INF])
    k = oracle.relative_kkt(lp, np.array([1.0, 3.0]), np.array([1.0, 0.0, 1.0]))
    assert k["rel"] <= 1e-15 and k["p"] == -7.0
    # a wrong-sign multiplier on a <= row is a dual-residual error
    assert oracle.relative_kkt(lp, np.array([1.0, 3.0]), np.array([-1.0, 0.0, 1.0]))["rel_d"] > 0.1


# ---- generators ------------------------------------------------------------

def test_generator_shapes_and_determinism(capi):
    a = capi.generate(0, seed=3)
    b = capi.generate(0, seed=3)
    c = capi.generate(0, seed=4)
    assert (a.m, a.n, a.nnz) == (2000, 5000, 40000)
    assert np.array_equal(a.values, b.values) and np.array_equal(a.c, b.c)
    assert not np.array_equal(a.values, c.values)
    assert np.all(a.values == a.values.astype(np.float32))
    assert np.count_nonzero(a.x_star) == 1500
    cols = np.bincount(a.col_idx, minlength=a.n)
    assert np.all(cols == 8)
    for i in range(0, a.m, 97):
        seg = a.col_idx[a.row_ptr[i]:a.row_ptr[i + 1]]
        assert np.all(np.diff(seg) > 0)


def test_generator_mixed_rows(capi, oracle):
    lp = capi.generate(None, seed=9, m=3000, n=6000, k_min=12, k_max=12, eq_frac=0.6, noise=0.05)
    ineq = lp.row_lower <= -INF
    assert abs(ineq.mean() - 0.4) < 0.01
    assert np.all(lp.y_star[ineq] == 0.0)
    act = sp.csr_matrix((lp.values, lp.col_idx, lp.row_ptr), shape=(lp.m, lp.n)) @ lp.x_star
    assert np.all(lp.row_upper[ineq] - act[ineq] > 0)
    assert oracle.relative_kkt(lp, lp.x_star, -lp.y_star)["rel"] <= 1e-13


def test_batch_instance_sizes(capi):
    ms = []
    for s in range(1, 40):
        p = capi.generate(3, seed=s)
        assert 50 <= p.m <= 5000 and p.n == int(2.5 * p.m + 0.5)
        k = np.bincount(p.col_idx, minlength=p.n)
        assert k.min() >= min(4, p.m) and k.max() <= 16
        ms.append(p.m)
    assert min(ms) < 300 and max(ms) > 1000


# ---- MPS -----------------------------------------------------------------

def test_mps_afiro_shape(capi):
    from pramaan_amd.mps import read_mps
    lp = read_mps(GOLDEN / "afiro.mps")
    assert (lp.m, lp.n) == (27, 32)
    assert lp.nnz == 83
    ad = read_mps(GOLDEN / "adlittle.mps")
    assert (ad.m, ad.n) == (56, 97)
