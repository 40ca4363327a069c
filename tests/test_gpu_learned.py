"""Learned warm start of a resident batch (pm_batch_learned_warm_start):
the device-built graph, features and predictions must equal the host path
(pm_gnn_lp_features + pm_gnn_set_graph + pm_gnn_forward) value for value,
and the judged warm start must solve every LP."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host_starts(capi, net, lps, pc, pr):
    from pramaan_amd import warmstart as W
    gs = [W.graph_of(lp, pc, pr) for lp in lps]
    net.set_graph(W.batch_graphs(gs))
    xh, yh = net.forward()
    out, xo, yo = [], 0, 0
    for lp, g in zip(lps, gs):
        out.append((np.clip(xh[xo:xo + g["n"]].astype(np.float64) * g["s"], lp.var_lower, lp.var_upper),
                    yh[yo:yo + g["m"]].astype(np.float64) * g["cs"]))
        xo += g["n"]
        yo += g["m"]
    return out


@pytest.mark.parametrize("kind", ["rolling", "generated"])
def test_device_warm_start_equals_host_path(capi, kind):
    if kind == "rolling":
        rh = capi.RollingHorizon()
        lps = [rh.window(w) for w in (0, 1, 5, 9)]
        pc, pr = rh.cols_per_period, rh.rows_per_period
    else:
        lps = [capi.generate(None, seed=s, m=40 + 30 * s, n=100 + 70 * s, eq_frac=0.6) for s in range(5)]
        pc = pr = 0
    net = capi.Gnn(seed=3)
    want = host_starts(capi, net, lps, pc, pr)
    with capi.Batch(lps, tolerance=1e-6, check_frequency=64) as b:
        secs = b.learned_warm_start(net, pc, pr)
        assert secs > 0
        for i, (wx, wy) in enumerate(want):
            x, y = b.get_warm_start(i)
            assert np.array_equal(x, wx), i
            assert np.array_equal(y, wy), i
        res, _, _ = b.run()
        acc, kkt = b.warm_start_verdicts()
    assert all(r.status == "OPTIMAL" for r in res)
    assert set(acc.tolist()) <= {0, 1} and np.all(kkt > 0)


def test_training_needs_a_training_graph(capi):
    lps = [capi.generate(None, seed=2, m=30, n=70)]
    net = capi.Gnn(seed=1)
    with capi.Batch(lps) as b:
        b.learned_warm_start(net)
    with pytest.raises(capi.PramaanError):
        net.train_step(np.zeros(70, np.float32), np.zeros(30, np.float32), lr=1e-3)


def test_trained_network_gains_over_cold_and_untrained(capi):
    """What the trained network (pramaan_amd/data/warmstart_gnn.npz) gains on 42
    windows spread over the rolling-horizon sequence, solved to 1e-6 in one
    batch: fewer iterations than the cold start, its predictions kept by the
    KKT judge, while an untrained network of the same shape gains nothing
    (its predictions are worse than the cold start and are replaced).
    Measured on 5,160 x 9,600 windows: 178,368 cold iterations, 160,192
    learned (-10 %, all 42 kept); the untrained network's starts are all
    replaced."""
    from pramaan_amd import rolling as R
    rh = capi.RollingHorizon()
    lps = [rh.window(w) for w in range(0, 500, 12)]
    cpp, rpp = rh.cols_per_period, rh.rows_per_period
    trained = R.load(R.default_weights())
    untrained = capi.Gnn(seed=7)
    out = {}
    with capi.Batch(lps, tolerance=1e-6, max_iterations=100000, check_frequency=64) as b:
        for name, net in (("cold", None), ("trained", trained), ("untrained", untrained)):
            for i in range(len(lps)):
                b.set_warm_start(i, None, None)
            if net is not None:
                b.learned_warm_start(net, cpp, rpp)
            res, _, _ = b.run()
            acc, _ = b.warm_start_verdicts()
            assert all(r.status == "OPTIMAL" for r in res), name
            out[name] = (sum(r.iterations for r in res), int((acc == 1).sum()))
    cold, trained_run, untrained_run = out["cold"][0], out["trained"], out["untrained"]
    assert trained_run[0] <= 0.95 * cold and trained_run[1] >= 40, out
    assert untrained_run[1] <= 4 and untrained_run[0] >= 0.95 * cold, out
