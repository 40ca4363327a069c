"""Infeasible and unbounded LPs: the reference PDHG has no infeasibility
or unboundedness detection, so it ends such a solve at its iteration limit
(src/gpu/pdhg_solver.cu:175, 392: status starts as kIterationLimit and only
the convergence test sets kOptimal).  Every path here must give that same
verdict: the restated reference on the CPU (the oracle), the reference
algorithm on the GPU, the PDLP algorithm, and the batch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF = 1e30


def cases(capi):
    return {
        # x1 + x2 <= -1 with x >= 0
        "primal_infeasible_row": capi.LP.from_dense([[1.0, 1.0]], [1.0, 1.0], [-INF], [-1.0], [0.0, 0.0],
                                                    [INF, INF]),
        # x1 + x2 = 1 and x1 + x2 = 2
        "primal_infeasible_equalities": capi.LP.from_dense([[1.0, 1.0], [1.0, 1.0]], [1.0, 2.0], [1.0, 2.0],
                                                           [1.0, 2.0], [0.0, 0.0], [INF, INF]),
        # min -x1 s.t. x1 - x2 <= 1, x >= 0: x1 = 1 + x2 grows without bound
        "unbounded": capi.LP.from_dense([[1.0, -1.0]], [-1.0, 0.0], [-INF], [1.0], [0.0, 0.0], [INF, INF]),
        # max x1 + x2 s.t. x1 - x2 = 0, x >= 0
        "unbounded_max": capi.LP.from_dense([[1.0, -1.0]], [1.0, 1.0], [0.0], [0.0], [0.0, 0.0], [INF, INF],
                                            sense=1),
    }


@pytest.mark.parametrize("name", ["primal_infeasible_row", "primal_infeasible_equalities", "unbounded",
                                  "unbounded_max"])
def test_verdict_matches_reference(capi, oracle, name):
    lp = cases(capi)[name]
    it = 3000
    ref, xr, _ = oracle.pdhg_solve(lp, tol=1e-6, max_iterations=it)
    assert ref["status"] == "ITERATION_LIMIT" and ref["iterations"] == it
    r = capi.solve(lp, tolerance=1e-6, max_iterations=it, spmv_variant="sequential")
    assert r.status == ref["status"] and r.iterations == ref["iterations"]
    assert np.array_equal(r.x, xr.astype(np.float64))
    p = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-6, max_iterations=it,
                   check_frequency=64)
    assert p.status == ref["status"] and p.iterations == it, (p.status, p.iterations, p.rel_kkt)
    with capi.Batch([lp, lp], tolerance=1e-6, max_iterations=it, check_frequency=64) as b:
        res, _, _ = b.run()
    for q in res:
        assert q.status == ref["status"] and q.iterations == it, (q.status, q.iterations, q.rel_kkt)
