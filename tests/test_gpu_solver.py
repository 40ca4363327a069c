"""GPU PDHG solver parity through the C-ABI (pm_pdhg_solve).

* Reference algorithm, sequential SpMV variant: the whole trajectory must be
  identical to the CPU oracle's restatement of PdhgSolver::solve
  (status, iteration count, restarts, pv, dv, objective, x bit for bit).
* Reference algorithm, fast variants: the known-answer LPs of
  tests/test_gpu_pdhg.cpp:263-404 and afiro (:409-452).
* PDLP algorithm (reflected Halpern + FP64 refinement): relative KKT and the
  objective against the closed-form optimum of the synthetic generator.
"""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
INF = 1e30


def known_lps(capi):
    D = capi.LP.from_dense
    return {
        "wyndor": (D([[1, 0], [0, 2], [3, 2]], [3, 5], [-INF] * 3, [4, 12, 18], [0, 0], [INF, INF], sense=1), 36.0),
        "zero_row": (capi.LP(np.array([2.0, -1.0]), np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0),
                             np.zeros(0), np.zeros(0), np.array([0.0, -5.0]), np.array([10.0, 5.0]), 0, 10.0), 5.0),
        "one_var": (D([[1.0]], [1.0], [3.0], [INF], [0.0], [INF]), 3.0),
        "equality": (D([[1, 1]], [1, 1], [5.0], [5.0], [0, 0], [INF, INF]), 5.0),
        "multi": (D([[1, 1], [1, 0], [0, 1]], [-1, -2], [-INF] * 3, [4, 3, 3], [0, 0], [INF, INF]), -7.0),
        "zero_obj": (D([[1, 1]], [0, 0], [2.0], [INF], [0, 0], [INF, INF]), 0.0),
        "bounded": (D([[1, 1]], [1, 1], [-INF], [10.0], [2, 3], [5, 7]), 5.0),
    }


def primal_violation(lp, x):
    v = max(np.max(lp.var_lower - x, initial=0.0), np.max(x - lp.var_upper, initial=0.0))
    act = np.array([np.dot(lp.values[lp.row_ptr[i]:lp.row_ptr[i + 1]], x[lp.col_idx[lp.row_ptr[i]:lp.row_ptr[i + 1]]])
                    for i in range(lp.m)])
    if lp.m:
        v = max(v, np.max(lp.row_lower - act, initial=0.0), np.max(act - lp.row_upper, initial=0.0))
    return v


@pytest.mark.parametrize("name", ["wyndor", "zero_row", "one_var", "equality", "multi", "zero_obj", "bounded"])
def test_reference_sequential_trajectory_matches_oracle(capi, oracle, name):
    lp, _ = known_lps(capi)[name]
    ref, xr, yr = oracle.pdhg_solve(lp)
    r = capi.solve(lp, spmv_variant="sequential")
    assert r.status == ref["status"]
    assert r.iterations == ref["iterations"]
    assert r.restarts == ref["restarts"]
    assert np.array_equal(r.x, xr.astype(np.float64))
    assert np.array_equal(r.y, yr.astype(np.float64))
    assert r.primal_residual == ref["primal_residual"]
    assert r.dual_residual == ref["dual_residual"]
    assert r.objective == pytest.approx(ref["objective"], rel=1e-15, abs=1e-15)


@pytest.mark.parametrize("name", ["wyndor", "zero_row", "one_var", "equality", "multi", "zero_obj", "bounded"])
def test_reference_known_answers(capi, name):
    lp, want = known_lps(capi)[name]
    r = capi.solve(lp)
    assert r.status == "OPTIMAL"
    assert abs(r.objective - want) <= 1e-3 + 1e-4 * abs(want)
    if lp.m:
        assert primal_violation(lp, r.x) <= 1e-4


def test_scaled_coefficients_terminate(capi):
    lp = capi.LP.from_dense([[1e3, 1e-3]], [1e-3, 1e3], [1.0], [INF], [0, 0], [INF, INF])
    r = capi.solve(lp)
    assert r.status in ("OPTIMAL", "ITERATION_LIMIT")
    if r.status == "OPTIMAL":
        assert primal_violation(lp, r.x) <= 1e-3


def test_nan_cost_is_numerical_failure(capi):
    lp = capi.LP.from_dense([[1.0]], [np.nan], [0.0], [1.0], [0.0], [1.0])
    assert capi.solve(lp).status == "NUMERICAL_FAILURE"


def test_reference_sequential_config0_matches_oracle(capi, oracle):
    """2,000 x 5,000 synthetic LP, reference algorithm, tol 1e-4: identical
    trajectory on the sequential variant."""
    lp = capi.generate(0, seed=1)
    ref, xr, _ = oracle.pdhg_solve(lp, tol=1e-4, max_iterations=20000)
    r = capi.solve(lp, tolerance=1e-4, max_iterations=20000, spmv_variant="sequential")
    assert r.status == ref["status"] == "OPTIMAL"
    assert r.iterations == ref["iterations"]
    assert np.array_equal(r.x, xr.astype(np.float64))
    assert r.objective == pytest.approx(ref["objective"], rel=1e-14)


def test_reference_fast_config0_objective(capi, oracle):
    """Fast SpMV variants change FP32 rounding; the objective at the same
    tolerance stays within the tolerance band of the oracle's and of p*."""
    lp = capi.generate(0, seed=1)
    ref, _, _ = oracle.pdhg_solve(lp, tol=1e-4, max_iterations=20000)
    r = capi.solve(lp, tolerance=1e-4, max_iterations=20000)
    assert r.status == "OPTIMAL"
    assert abs(r.objective - ref["objective"]) <= 1e-3 * abs(ref["objective"])
    assert abs(r.objective - lp.p_star) <= 1e-3 * abs(lp.p_star)
    assert r.primal_residual <= 1e-4 and r.dual_residual <= 1e-4


def test_afiro_reference_algorithm(capi):
    from pramaan_amd.mps import read_mps
    lp = read_mps(GOLDEN / "afiro.mps")
    r = capi.solve(lp, max_iterations=200000, check_frequency=500, tolerance=1e-5)
    assert r.status == "OPTIMAL"
    assert abs(r.objective - (-464.75314286)) <= 1e-4 * 10  # reference: GPU vs CPU within 1e-4 abs
    assert primal_violation(lp, r.x) <= 1e-4


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pdlp_config0_to_1e8(capi, oracle, seed):
    lp = capi.generate(0, seed=seed)
    r = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-8, max_iterations=100000,
                   check_frequency=64)
    assert r.status == "OPTIMAL", (r.iterations, r.rel_kkt)
    k = oracle.relative_kkt(lp, r.x, r.y)
    assert k["rel"] <= 1.5e-8
    assert abs(k["rel"] - r.rel_kkt) <= 1e-12 + 1e-6 * r.rel_kkt
    assert abs(r.objective - lp.p_star) <= 1e-8 * (1 + abs(lp.p_star))


def test_pdlp_afiro(capi, oracle):
    from pramaan_amd.mps import read_mps
    lp = read_mps(GOLDEN / "afiro.mps")
    r = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-8, max_iterations=200000,
                   check_frequency=64)
    assert r.status == "OPTIMAL"
    assert r.objective == pytest.approx(-464.75314286, abs=1e-5)


def test_pdlp_warm_start_from_optimum(capi):
    lp = capi.generate(0, seed=4)
    cold = capi.solve(lp, algorithm="pdlp", tolerance=1e-7, check_frequency=64)
    warm = capi.solve(lp, algorithm="pdlp", tolerance=1e-7, check_frequency=64, x0=cold.x, y0=cold.y)
    assert cold.status == warm.status == "OPTIMAL"
    assert warm.iterations <= cold.iterations


@pytest.mark.parametrize("which", ["generated", "afiro", "dense_rows"])
def test_device_scaling_matches_restatement(capi, which, ruiz_restated):
    """The device Ruiz + Pock-Chambolle factors equal a plain-loop
    restatement bit for bit (row and column norms in the same order; rows
    of more than 256 entries, here 3 rows of 750, are reduced by a whole
    wavefront in the maximum passes)."""
    if which == "afiro":
        from pramaan_amd.mps import read_mps
        lp = read_mps(GOLDEN / "afiro.mps")
    elif which == "dense_rows":
        lp = capi.with_dense_rows(capi.generate(None, seed=21, m=300, n=750, eq_frac=0.6, noise=0.05), 3)
    else:
        lp = capi.generate(None, seed=21, m=300, n=750, eq_frac=0.6, noise=0.05)
    ctx = capi.Context(0)
    ctx.load(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-6)
    dr, dc = ctx.scaling()
    ctx.close()
    wr, wc = ruiz_restated(lp)
    assert np.array_equal(dr, wr) and np.array_equal(dc, wc), (np.abs(dr - wr).max(), np.abs(dc - wc).max())


@pytest.mark.parametrize("case", ["config0", "window", "afiro", "dense_rows", "merge"])
def test_graph_chunks_match_direct_launches(capi, case):
    """Full chunks of check_frequency iterations replayed as one captured
    graph (epilogues reading tau, sigma, the dual shift's presence and the
    Halpern weights from device memory) give the same bits as the same
    launches issued directly: refinement rounds, restarts and (on the
    window) feasibility polishing included."""
    kw = dict(algorithm="pdlp", termination="relative_kkt", tolerance=1e-6, max_iterations=200000,
              check_frequency=64)
    if case == "config0":
        lp, kw["tolerance"] = capi.generate(0, seed=1), 1e-8
    elif case == "window":
        lp = capi.RollingHorizon().window(166)
    elif case == "afiro":
        from pramaan_amd.mps import read_mps
        lp, kw["tolerance"] = read_mps(GOLDEN / "afiro.mps"), 1e-8
    elif case == "dense_rows":
        lp, kw["max_iterations"] = capi.with_dense_rows(capi.generate(2, seed=1), 16), 1000
    else:
        lp, kw["spmv_variant"] = capi.generate(0, seed=2), "merge"
    out = {}
    for use_graph in (True, False):
        with capi.Context(0) as ctx:
            ctx.set_graph(use_graph)
            ctx.load(lp, **kw)
            out[use_graph] = ctx.solve()
    a, b = out[True], out[False]
    print(f"{case}: {a.status} it={a.iterations} graph {a.device_seconds:.4f}s direct {b.device_seconds:.4f}s",
          flush=True)
    assert a.status == b.status and a.iterations == b.iterations and a.restarts == b.restarts
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
