"""A PDLP solve that stops at its iteration limit returns the checked
candidate of lowest relative KKT error.  Solves are deterministic, so a
solve with a larger limit passes through every check of one with a smaller
limit: its returned error can only be lower or equal.  On the 5,000 x
10,000 LP of the single-large-LP generator's shape (60 % equality rows, 5 %
value noise) the error rises again late in a stalled solve: returning the
last candidate gave 1.4e-6 after 80,000 iterations and 8.4e-6 after
200,000, the best candidates are 1.2e-6 and 9.6e-8.  The oracle's FP64
relative KKT of the returned (x, y) must equal the reported one."""
import pytest

pytestmark = pytest.mark.gpu


def test_limit_returns_best_checked_candidate(capi, oracle):
    lp = capi.generate(None, seed=1, m=5000, n=10000, k_min=12, k_max=12, eq_frac=0.6, noise=0.05, support=0.3,
                       var_upper=10.0)
    rels = []
    for limit in (20000, 80000, 200000):
        r = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-10, max_iterations=limit,
                       check_frequency=64)
        k = oracle.relative_kkt(lp, r.x, r.y)
        print(f"limit {limit}: {r.status} rel_kkt {r.rel_kkt:.3e} (oracle {k['rel']:.3e})", flush=True)
        assert r.status == "ITERATION_LIMIT" and r.iterations == limit
        assert abs(k["rel"] - r.rel_kkt) <= 1e-6 * r.rel_kkt
        rels.append(r.rel_kkt)
    assert rels[1] <= rels[0] and rels[2] <= rels[1], rels
