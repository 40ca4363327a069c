"""Warm-start network on the GPU (pm_gnn_*) against the CPU oracle
(oracle/gnn_oracle.py, FP64 with the same bfloat16 rounding of every 64x64
product's inputs).

Tolerances: FP32 accumulation and the occasional different bf16 rounding of
an activation that sits on a rounding boundary make per-entry differences
of a few bf16 ulps possible, so outputs are compared at 5e-3 relative to
their scale and gradients per tensor by relative 2-norm error <= 2e-2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_problem(capi, seed=3, m=60, n=150):
    from pramaan_amd import warmstart as W
    lp = capi.generate(None, seed=seed, m=m, n=n, k_min=3, k_max=6, eq_frac=0.7)
    g = W.graph_of(lp, None, None)
    rng = np.random.default_rng(seed)
    tx = rng.uniform(0, 1, n)
    ty = rng.standard_normal(m)
    return lp, g, tx, ty


def oracle_graph(go, g):
    return go.Graph(g["m"], g["n"], g["rp"], g["ci"], g["ar"].astype(np.float32).astype(np.float64),
                    g["cp"], g["ri"], g["ac"].astype(np.float32).astype(np.float64))


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def test_forward_matches_oracle(capi):
    from oracle import gnn_oracle as go
    _, g, _, _ = small_problem(capi)
    net = capi.Gnn(seed=5)
    net.set_graph(g)
    x, y = net.forward()
    p = go.unflatten(net.get_params().astype(np.float64))
    c = go.forward(p, oracle_graph(go, g), f32(g["xc"]), f32(g["xr"]), use_bf16=True)
    sx = np.abs(c["xhat"]).max() + 1e-12
    sy = np.abs(c["yhat"]).max() + 1e-12
    assert np.max(np.abs(x - c["xhat"])) <= 5e-3 * sx
    assert np.max(np.abs(y - c["yhat"])) <= 5e-3 * sy


def test_gradient_matches_oracle(capi):
    from oracle import gnn_oracle as go
    _, g, tx, ty = small_problem(capi, seed=4)
    net = capi.Gnn(seed=6)
    net.set_graph(g)
    loss, grad = net.gradient(tx, ty)
    p = go.unflatten(net.get_params().astype(np.float64))
    oloss, ogr, _ = go.loss_and_grad(p, oracle_graph(go, g), f32(g["xc"]), f32(g["xr"]), f32(tx), f32(ty),
                                     use_bf16=True)
    assert loss == pytest.approx(oloss, rel=5e-3)
    gg = go.unflatten(grad.astype(np.float64))
    for name, _ in go.param_shapes():
        a, b = gg[name].ravel(), ogr[name].ravel()
        den = np.linalg.norm(b)
        if den < 1e-9:
            assert np.linalg.norm(a) < 1e-6, name
            continue
        assert np.linalg.norm(a - b) <= 2e-2 * den, (name, np.linalg.norm(a - b) / den)


def test_primal_weight_head_matches_oracle(capi):
    """Three LPs in one block-diagonal graph: the head's predictions and the
    gradient of the loss with its term, against the oracle."""
    from oracle import gnn_oracle as go
    from pramaan_amd import warmstart as W
    lps = [capi.generate(None, seed=s, m=30 + 10 * s, n=70 + 25 * s, k_min=3, k_max=6, eq_frac=0.7)
           for s in range(3)]
    gs = [W.graph_of(lp, None, None) for lp in lps]
    g = W.batch_graphs(gs)
    seg = np.concatenate([[0], np.cumsum([x["n"] for x in gs])])
    rng = np.random.default_rng(12)
    tx, ty = rng.uniform(0, 1, g["n"]), rng.standard_normal(g["m"])
    tw = np.array([0.4, -1.1, 2.0])
    net = capi.Gnn(seed=13)
    p0 = net.get_params()
    names = [name for name, shp in go.param_shapes() for _ in range(int(np.prod(shp)))]
    p0[[k for k, nm in enumerate(names) if nm == "wom"]] = rng.standard_normal(64) * 0.2  # a non-zero head
    net.set_params(p0)
    net.set_graph(g)
    net.set_segments(seg)
    what = net.forward_omega()
    p = go.unflatten(net.get_params().astype(np.float64))
    og = oracle_graph(go, g)
    c = go.forward(p, og, f32(g["xc"]), f32(g["xr"]), use_bf16=True, seg=seg)
    assert np.max(np.abs(what - c["what"])) <= 5e-3 * (np.abs(c["what"]).max() + 1e-6)
    net.set_omega_targets(tw)
    loss, grad = net.gradient(tx, ty)
    oloss, ogr, _ = go.loss_and_grad(p, og, f32(g["xc"]), f32(g["xr"]), f32(tx), f32(ty), use_bf16=True, seg=seg,
                                     tw=tw)
    assert loss == pytest.approx(oloss, rel=5e-3)
    gg = go.unflatten(grad.astype(np.float64))
    for name, _ in go.param_shapes():
        den = np.linalg.norm(ogr[name])
        if den > 1e-9:
            assert np.linalg.norm(gg[name] - ogr[name]) <= 2e-2 * den, name


def test_adam_step_matches_oracle(capi):
    from oracle import gnn_oracle as go
    _, g, tx, ty = small_problem(capi, seed=7)
    net = capi.Gnn(seed=8)
    net.set_graph(g)
    p0 = net.get_params().astype(np.float64)
    _, grad = net.gradient(tx, ty)
    net.train_step(tx, ty, lr=1e-3)
    p1 = net.get_params().astype(np.float64)
    want, _, _ = go.adam(p0, grad.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, lr=1e-3)
    assert np.max(np.abs(p1 - want)) <= 1e-6


def test_training_reduces_loss(capi):
    _, g, tx, ty = small_problem(capi, seed=9)
    net = capi.Gnn(seed=10)
    net.set_graph(g)
    first = net.train_step(tx, ty, lr=3e-3)
    for _ in range(200):
        last = net.train_step(tx, ty, lr=3e-3)
    assert last < 0.5 * first


def test_bad_graphs_are_rejected(capi):
    """pm_gnn_set_graph checks the graph on the device while it builds the
    transposed value orders: an index out of range, a row whose columns do
    not increase, or a column order that is not the transpose of the row
    order fails with PM_ERR_INVALID_ARG and clears the graph; a good graph
    loads again afterwards."""
    _, g, tx, ty = small_problem(capi, seed=7)
    net = capi.Gnn(seed=8)
    net.set_graph(g)
    good, _ = net.gradient(tx, ty)
    bad = dict(g)
    bad["ci"] = g["ci"].copy()
    bad["ci"][5] = g["n"] + 3
    swapped = dict(g)                     # two columns of a row out of order
    swapped["ci"] = g["ci"].copy()
    r = next(i for i in range(g["m"]) if g["rp"][i + 1] - g["rp"][i] >= 2)
    a = int(g["rp"][r])
    swapped["ci"][a], swapped["ci"][a + 1] = swapped["ci"][a + 1], swapped["ci"][a]
    moved = dict(g)                       # a column entry that the rows do not hold
    moved["ri"] = g["ri"].copy()
    c = next(j for j in range(g["n"]) if g["cp"][j + 1] - g["cp"][j] >= 1)
    q = int(g["cp"][c])
    moved["ri"][q] = (int(g["ri"][q]) + 1) % g["m"]
    for case in (bad, swapped, moved):
        with pytest.raises(capi.PramaanError) as e:
            net.set_graph(case)
        assert e.value.code == 1
    net.set_graph(g)
    again, _ = net.gradient(tx, ty)
    assert again == good
