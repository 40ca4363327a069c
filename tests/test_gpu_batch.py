"""Batched solver (pm_batch_*): one launch, one workgroup per LP.

Every instance is checked against its own closed-form optimum and against
the oracle's FP64 relative KKT of the returned point; the batch must agree
with the single-LP engine, which runs the same decisions (pm::Control).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_batch(capi, count=24, seed0=100):
    lps = []
    for s in range(count):
        m = int(np.exp(np.random.default_rng(seed0 + s).uniform(np.log(30), np.log(800))))
        lps.append(capi.generate(None, seed=seed0 + s, m=m, n=int(2.5 * m + 0.5), k_min=min(4, m),
                                 k_max=min(16, m), eq_frac=1.0 if s % 2 else 0.7))
    return lps


def test_batch_matches_closed_form(capi, oracle):
    lps = small_batch(capi)
    with capi.Batch(lps, tolerance=1e-7, max_iterations=100000, check_frequency=64) as b:
        res, secs, inst = b.run(want_vectors=True)
    assert secs > 0 and np.all(inst > 0)
    for lp, r in zip(lps, res):
        assert r.status == "OPTIMAL", (lp.m, r.iterations, r.rel_kkt)
        k = oracle.relative_kkt(lp, r.x, r.y)
        assert k["rel"] <= 1.2e-7
        assert abs(k["rel"] - r.rel_kkt) <= 1e-12 + 1e-6 * r.rel_kkt
        assert abs(r.objective - lp.p_star) <= 1e-6 * (1 + abs(lp.p_star))


def test_batch_replay_is_deterministic(capi):
    lps = small_batch(capi, count=8, seed0=300)
    with capi.Batch(lps, tolerance=1e-6, check_frequency=64) as b:
        r1, _, _ = b.run(want_vectors=True)
        r2, _, _ = b.run(want_vectors=True)
    for a, c in zip(r1, r2):
        assert a.iterations == c.iterations and np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)


def test_batch_agrees_with_single_lp_engine(capi):
    lps = small_batch(capi, count=4, seed0=500)
    with capi.Batch(lps, tolerance=1e-8, check_frequency=64) as b:
        rb, _, _ = b.run(want_vectors=True)
    for lp, r in zip(lps, rb):
        s = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-8, check_frequency=64)
        assert s.status == r.status == "OPTIMAL"
        assert abs(s.objective - r.objective) <= 1e-8 * (1 + abs(s.objective))


def test_batch_warm_start_from_solution(capi):
    lps = small_batch(capi, count=6, seed0=700)
    with capi.Batch(lps, tolerance=1e-7, check_frequency=64) as b:
        cold, _, _ = b.run(want_vectors=True)
        for i, r in enumerate(cold):
            b.set_warm_start(i, r.x, r.y)
        warm, _, _ = b.run(want_vectors=True)
    for c, w in zip(cold, warm):
        assert w.status == "OPTIMAL" and w.iterations <= c.iterations


def test_batch_rejects_bad_input(capi):
    lp = capi.generate(None, seed=1, m=20, n=50, k_min=4, k_max=4)
    lp.col_idx[0] = 10**6
    with pytest.raises(capi.PramaanError):
        capi.Batch([lp])
