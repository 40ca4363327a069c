"""Batched solver (pm_batch_*): many LPs solved together in one graph-captured run.

Every instance is checked against its own closed-form optimum and against
the oracle's FP64 relative KKT of the returned point; the batch must agree
with the single-LP engine, which runs the same decisions (pm::Control).
"""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"

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
        # A relative KKT of 1e-7 leaves primal residuals of ~1e-7 (1 + |b|);
        # priced by |y*| they move the objective by up to ~1e-5.
        assert abs(r.objective - lp.p_star) <= 1e-5 * (1 + abs(lp.p_star))


def test_batch_replay_is_deterministic(capi):
    lps = small_batch(capi, count=8, seed0=300)
    with capi.Batch(lps, tolerance=1e-6, check_frequency=64) as b:
        r1, _, _ = b.run(want_vectors=True)
        r2, _, _ = b.run(want_vectors=True)
    for a, c in zip(r1, r2):
        assert a.iterations == c.iterations and np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)


def test_batch_direct_launches_match_graph(capi):
    """The captured graph and the same launches issued directly give the
    same bits (every reduction in the engine has a fixed order)."""
    lps = small_batch(capi, count=6, seed0=900)
    with capi.Batch(lps, tolerance=1e-6, check_frequency=32) as b:
        rg, _, _ = b.run(want_vectors=True)
        b.set_graph(False)
        rd, _, _ = b.run(want_vectors=True)
    for a, c in zip(rg, rd):
        assert a.status == c.status == "OPTIMAL"
        assert a.iterations == c.iterations and np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)


def test_batch_sliced_kernels_match_one_lane_per_row(capi):
    """The sliced kernels (global and LDS gathers) sum every row in CSR
    order with FMA, as the one-lane-per-row kernel does, so whole solves
    give the same bits; the batch mixes full groups (2048 rows of A for the
    2300-row LP, whose random columns make its groups wide; 2048 of A^T;
    walked in passes of 16 slices) with partial groups and slices, an
    empty row and a 1500-entry row."""
    lps = small_batch(capi, count=5, seed0=1700)
    lps.append(capi.generate(None, seed=1705, m=2300, n=5750, k_min=4, k_max=16))
    nd = 1500
    dense = np.vstack([np.ones(nd), np.eye(nd)[:3], np.zeros(nd)])
    lps.append(capi.LP.from_dense(dense, np.linspace(1.0, 2.0, nd), [10.0, 0.0, 0.0, 0.0, -1.0],
                                  [10.0, 5.0, 5.0, 5.0, 1.0], np.zeros(nd), np.full(nd, 1e30)))
    out = {}
    for variant in ("scalar", "sliced", "staged"):
        with capi.Batch(lps, tolerance=1e-7, max_iterations=60000, check_frequency=64, spmv_variant=variant) as b:
            out[variant], _, _ = b.run(want_vectors=True)
    for variant in ("sliced", "staged"):
        for a, c in zip(out["scalar"], out[variant]):
            assert a.status == c.status == "OPTIMAL"
            assert a.iterations == c.iterations and a.objective == c.objective, variant
            assert np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y), variant


def test_batch_size_class_launches_match_one_launch(capi):
    """One iterate launch per size class (m < 256, m < 1024, larger) and one
    launch over every LP run the same kernel on the same work items, so a
    batch with LPs of all three classes gives the same bits either way."""
    lps = small_batch(capi, count=10, seed0=2100)
    lps.append(capi.generate(None, seed=2111, m=1500, n=3750, k_min=4, k_max=16))
    out = {}
    for launches in (1, 3):
        with capi.Batch(lps, tolerance=1e-7, max_iterations=60000, check_frequency=64, spmv_variant="staged") as b:
            b.set_size_classes(launches)
            info = b.size_classes()
            out[launches], _, _ = b.run(want_vectors=True)
        assert info["launches"] == launches
    assert all(k > 0 for k in info["lps"]), info
    for a, c in zip(out[1], out[3]):
        assert a.status == c.status == "OPTIMAL"
        assert a.iterations == c.iterations and np.array_equal(a.x, c.x) and np.array_equal(a.y, c.y)


def test_batch_device_scaling_matches_restatement(capi, ruiz_restated):
    """The batch scales its block-diagonal concatenation on the device; each
    LP's factors equal the plain-loop restatement of that LP alone, bit for
    bit (an LP without rows, an LP with an empty row and column, mixed
    sizes, rows of 500 entries reduced by a whole wavefront)."""
    from pramaan_amd.mps import read_mps
    lps = [capi.generate(None, seed=31, m=120, n=300, eq_frac=0.6, noise=0.05),
           read_mps(GOLDEN / "afiro.mps"),
           capi.LP.from_dense(np.zeros((0, 3)), [1.0, -1.0, 2.0], [], [], np.zeros(3), np.ones(3)),
           capi.LP.from_dense([[2.0, 0.0, -4.0], [0.0, 0.0, 0.0], [0.5, 0.0, 8.0]], [1.0, 0.0, 1.0],
                              [1.0, -1.0, 0.0], [4.0, 1.0, 9.0], np.zeros(3), np.full(3, 5.0)),
           capi.generate(None, seed=32, m=700, n=1750, k_min=4, k_max=16),
           capi.with_dense_rows(capi.generate(None, seed=33, m=200, n=500, k_min=4, k_max=16), 2)]
    with capi.Batch(lps, tolerance=1e-6, check_frequency=64) as b:
        for i, lp in enumerate(lps):
            dr, dc = b.scaling(i)
            wr, wc = ruiz_restated(lp)
            assert np.array_equal(dr, wr) and np.array_equal(dc, wc), i


def test_batch_wide_window_falls_back(capi):
    """A row touching 20,000 columns makes A's column window wider than the
    staged kernel's LDS holds: an explicit PM_SPMV_STAGED is rejected at
    creation, and the autotuner picks another kernel for A and still solves
    the batch (min c.x, sum x = 10, x_0..x_2 <= 5: x_0 = x_1 = 5)."""
    nd = 20000
    dense = np.vstack([np.ones(nd), np.eye(nd)[:3]])
    c = np.linspace(1.0, 2.0, nd)
    lp = capi.LP.from_dense(dense, c, [10.0, 0.0, 0.0, 0.0], [10.0, 5.0, 5.0, 5.0], np.zeros(nd),
                            np.full(nd, 1e30))
    lps = small_batch(capi, count=3, seed0=1900) + [lp]
    with pytest.raises(capi.PramaanError, match="window"):
        capi.Batch(lps, tolerance=1e-6, spmv_variant="staged")
    with capi.Batch(lps, tolerance=1e-7, max_iterations=60000, check_frequency=64) as b:
        va, _ = b.spmv_variants()
        assert va != capi.SPMV["staged"]
        res, _, _ = b.run()
    for q, r in zip(lps[:-1], res):
        assert r.status == "OPTIMAL" and abs(r.objective - q.p_star) <= 1e-5 * (1 + abs(q.p_star))
    assert res[-1].status == "OPTIMAL" and abs(res[-1].objective - 5 * (c[0] + c[1])) <= 1e-4


def test_batch_one_lp_and_mixed_sizes(capi):
    """A batch of one LP and a batch whose sizes differ 100x (the work list
    and the per-LP chunk lengths), against each LP's closed form."""
    tiny = capi.generate(None, seed=11, m=5, n=12, k_min=1, k_max=3)
    big = capi.generate(None, seed=12, m=600, n=1500, k_min=4, k_max=16)
    for lps in ([big], [tiny, big, tiny], [big, tiny]):
        with capi.Batch(lps, tolerance=1e-7, max_iterations=50000, check_frequency=64) as b:
            res, _, _ = b.run()
        for lp, r in zip(lps, res):
            assert r.status == "OPTIMAL", (lp.m, r.iterations, r.rel_kkt)
            assert abs(r.objective - lp.p_star) <= 1e-5 * (1 + abs(lp.p_star))


@pytest.mark.parametrize("variant", ["stream", "sliced", "staged", "scalar", "vector4", "vector32"])
def test_batch_iterate_kernels_agree(capi, variant):
    """Every iterate kernel (LDS-staged stream segments, sliced ELL, sliced ELL
    gathering from LDS, 1..32 lanes per row)
    solves the batch, including a row longer than one stream segment, to the
    closed-form optimum."""
    lps = small_batch(capi, count=6, seed0=1300)
    nd = 1500  # one row longer than a stream segment (1024 nonzeros)
    dense = capi.LP.from_dense(np.vstack([np.ones(nd), np.eye(nd)[:3]]), np.linspace(1.0, 2.0, nd),
                               [10.0, 0.0, 0.0, 0.0], [10.0, 5.0, 5.0, 5.0], np.zeros(nd), np.full(nd, 1e30))
    lps.append(dense)
    with capi.Batch(lps, tolerance=1e-7, max_iterations=60000, check_frequency=64, spmv_variant=variant) as b:
        assert b.spmv_variants() == (capi.SPMV[variant], capi.SPMV[variant])
        res, _, _ = b.run()
    for lp, r in zip(lps[:-1], res):
        assert r.status == "OPTIMAL", (variant, lp.m, r.iterations, r.rel_kkt)
        assert abs(r.objective - lp.p_star) <= 1e-5 * (1 + abs(lp.p_star))
    # min c.x, sum x = 10, x_0, x_1, x_2 <= 5, x >= 0, c increasing: x_0 = x_1 = 5
    c = np.linspace(1.0, 2.0, nd)
    assert res[-1].status == "OPTIMAL" and abs(res[-1].objective - 5 * (c[0] + c[1])) <= 1e-4


def test_batch_iteration_limit(capi):
    """max_iterations that is not a multiple of check_frequency: the last
    chunk is cut short and the LP reports the limit with that count."""
    lps = small_batch(capi, count=3, seed0=40)
    with capi.Batch(lps, tolerance=1e-12, max_iterations=100, check_frequency=64) as b:
        res, _, _ = b.run()
    for r in res:
        assert r.status == "ITERATION_LIMIT" and r.iterations == 100


def test_batch_kernel_timing(capi):
    lps = small_batch(capi, count=8, seed0=60)
    with capi.Batch(lps, tolerance=1e-6, check_frequency=64) as b:
        for which in (capi.PM_KERNEL_SPMV_A, capi.PM_KERNEL_SPMV_AT):
            ms, by = b.time_kernel(which, 5)
            assert ms > 0 and by > 8 * sum(lp.nnz for lp in lps)
        res, _, _ = b.run()  # a run after timing starts afresh
    assert all(r.status == "OPTIMAL" for r in res)


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


def test_primal_weight_stays_bounded(capi):
    """Generated LP seed 2007 of the 4,096-LP workload (3709 x 9273): with an
    unbounded primal-weight update the weight ran to 4.6e5 and the gap
    stalled at 1.0 for 200k iterations; bounded within 100x of its initial
    value, both engines solve it."""
    lp = capi.generate(3, seed=2007)
    with capi.Batch([lp], tolerance=1e-6, max_iterations=50000, check_frequency=64) as b:
        (r,), _, _ = b.run()
    assert r.status == "OPTIMAL", (r.iterations, r.rel_kkt, r.primal_weight)
    assert abs(r.objective - lp.p_star) <= 1e-4 * (1 + abs(lp.p_star))
    s = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-6, max_iterations=50000,
                   check_frequency=64)
    assert s.status == "OPTIMAL" and abs(s.objective - r.objective) <= 1e-5 * (1 + abs(r.objective))
