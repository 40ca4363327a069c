"""Feasibility polishing of the PDLP engines (pm_pdhg_options.
feasibility_polishing; Polish in pramaan_amd/csrc/control.hpp): once the
relative gap of a check is small, the primal and the dual feasibility
problems are solved from the candidate and the pair is accepted when its
relative KKT error meets the tolerance with a gap of at most half of it.

Checked here, on rolling-horizon windows of 5,160 x 9,600:
  * every window solved with polishing is OPTIMAL and its returned (x, y)
    has an oracle FP64 relative KKT error of at most 1.2e-6 (20 % head room
    for the FP64 summation order, as in the other window tests);
  * polishing is used (attempts are counted) and does not cost iterations
    over the same batch without it;
  * a polished solution's objective is within 1e-6 relative of HiGHS
    (scipy's linprog, computed here), the accuracy the acceptance rule's
    half-tolerance gap is there to keep;
  * with polishing off the engine reports no attempts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-6
KKT_TOL = 1.2e-6
OBJ_TOL = 1e-6


def highs_objective(lp):
    import scipy.sparse as sp
    from scipy.optimize import linprog
    A = sp.csr_matrix((lp.values, lp.col_idx, lp.row_ptr), shape=(lp.m, lp.n))
    big = 1e29
    lo, hi = lp.row_lower, lp.row_upper
    eq = np.where((lo > -big) & (hi < big) & (lo == hi))[0]
    up = np.setdiff1d(np.where(hi < big)[0], eq)
    dn = np.setdiff1d(np.where(lo > -big)[0], eq)
    bounds = np.stack([np.where(lp.var_lower > -big, lp.var_lower, -np.inf),
                       np.where(lp.var_upper < big, lp.var_upper, np.inf)], 1)
    r = linprog(lp.c, A_ub=sp.vstack([A[up], -A[dn]]), b_ub=np.concatenate([hi[up], -lo[dn]]), A_eq=A[eq],
                b_eq=lo[eq], bounds=bounds, method="highs")
    assert r.status == 0, r.message
    return float(r.fun) + lp.offset


def test_batch_windows_with_and_without_polishing(capi, oracle):
    rh = capi.RollingHorizon()
    wins = list(range(0, 500, 10))
    lps = [rh.window(w) for w in wins]
    out = {}
    for pol in (1, 0):
        with capi.Batch(lps, tolerance=TOL, max_iterations=200000, check_frequency=64,
                        feasibility_polishing=pol) as b:
            res, _, _ = b.run(want_vectors=True)
        for w, lp, r in zip(wins, lps, res):
            assert r.status == "OPTIMAL", (pol, w, r.status)
            k = oracle.relative_kkt(lp, r.x, r.y)
            assert k["rel"] <= KKT_TOL, (pol, w, k)
        out[pol] = res
    attempts = np.array([r.polish_attempts for r in out[1]])
    assert (attempts > 0).sum() >= 5, attempts
    assert all(r.polish_attempts == 0 for r in out[0])
    it_on = sum(r.iterations for r in out[1])
    it_off = sum(r.iterations for r in out[0])
    assert it_on <= it_off, (it_on, it_off)


def test_polished_objective_matches_highs(capi, oracle):
    """Windows whose single-LP solve polishes: objective against HiGHS."""
    rh = capi.RollingHorizon()
    polished = 0
    for w in (166, 465, 133):
        lp = rh.window(w)
        r = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=TOL, max_iterations=200000,
                       check_frequency=64)
        assert r.status == "OPTIMAL", (w, r.status)
        k = oracle.relative_kkt(lp, r.x, r.y)
        assert k["rel"] <= KKT_TOL, (w, k)
        want = highs_objective(lp)
        assert abs(r.objective - want) <= OBJ_TOL * (1 + abs(want)), (w, r.objective, want)
        polished += r.polish_attempts > 0
        off = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=TOL, max_iterations=200000,
                         check_frequency=64, feasibility_polishing=0)
        assert off.status == "OPTIMAL" and off.polish_attempts == 0, (w, off.status)
    assert polished >= 1
