"""The column-blocked A x-bar of the single-LP PDLP engine
(PM_SPMV_COLBLOCK): A in column blocks, one launch per block, row sums
accumulated across blocks.  Forced on the 50,000 x 100,000 LP of the
single-large-LP generator (below the size where AUTO tries it), the solve
must reach the same optimum as the plain row kernels: its objective within
1e-6 of the closed form and the oracle's FP64 relative KKT error of its
(x, y) within the tolerance.  The blocked sums add the blocks in a different
order than a row walk, so the two trajectories differ in rounding only."""
import pytest

pytestmark = pytest.mark.gpu


def test_colblock_solve_matches_plain(capi, oracle):
    lp = capi.generate(2, seed=1)
    tol = 1e-4
    kw = dict(algorithm="pdlp", termination="relative_kkt", tolerance=tol, max_iterations=60000, check_frequency=64)
    with capi.Context(0) as ctx:
        ctx.load(lp, spmv_variant="colblock", **kw)
        va, _ = ctx.spmv_variants()
        assert va == "colblock"
        rb = ctx.solve()
    rp = capi.solve(lp, spmv_variant="auto", **kw)
    print(f"colblock: {rb.status} it={rb.iterations} obj={rb.objective:.10g}; plain: {rp.status} it={rp.iterations} "
          f"obj={rp.objective:.10g}; p*={lp.p_star:.10g}", flush=True)
    for r in (rb, rp):
        assert r.status == "OPTIMAL"
        assert abs(r.objective - lp.p_star) <= 1e-6 * (1 + abs(lp.p_star))
        assert oracle.relative_kkt(lp, r.x, r.y)["rel"] <= 1.2 * tol


def test_colblock_full_size_tracks_row_kernels(capi):
    """The 500,000 x 1,000,000 LP (12M nonzeros, x-bar 4 MB): 640 iterations
    with the column blocks forced and with the 8-lane row kernel forced for
    A x-bar walk the same trajectory up to summation order.  The PDHG step
    is non-expansive, so rounding differences (FP32, about 6e-8 relative
    per product) add up at most linearly: 640 iterations allow about 4e-5;
    the bound is 2e-5 on the objective (measured 2.5e-6) and 1e-4 on the
    relative KKT error (measured 1.3e-8)."""
    lp = capi.generate(1, seed=1)
    kw = dict(algorithm="pdlp", termination="relative_kkt", tolerance=1e-9, max_iterations=640, check_frequency=64)
    out = {}
    for var in ("colblock", "vector8"):
        with capi.Context(0) as ctx:
            ctx.load(lp, spmv_variant=var, **kw)
            assert ctx.spmv_variants()[0] == var
            out[var] = ctx.solve(want_vectors=False)
    a, b = out["colblock"], out["vector8"]
    print(f"colblock obj={a.objective:.12g} kkt={a.rel_kkt:.6e}; vector8 obj={b.objective:.12g} kkt={b.rel_kkt:.6e}",
          flush=True)
    assert a.status == b.status == "ITERATION_LIMIT"
    assert abs(a.objective - b.objective) <= 2e-5 * (1 + abs(b.objective))
    assert abs(a.rel_kkt - b.rel_kkt) <= 1e-4 * b.rel_kkt
