"""Pins the GNN oracle (oracle/gnn_oracle.py) before it checks the GPU:
central finite differences on a small random bipartite graph (FP64, no
bf16 rounding), the bf16 rounding rule against exact cases, and Adam against
its closed form on one scalar."""
import numpy as np
import pytest


def tiny_graph(go, m=7, n=11, seed=0):
    rng = np.random.default_rng(seed)
    A = np.zeros((m, n))
    for j in range(n):
        for i in rng.choice(m, 2, replace=False):
            A[i, j] = rng.standard_normal()
    rp, ci, ar = [0], [], []
    for i in range(m):
        nz = np.nonzero(A[i])[0]
        mx = np.abs(A[i, nz]).max() if len(nz) else 1.0
        ci += list(nz)
        ar += list(A[i, nz] / mx)
        rp.append(len(ci))
    cp, ri, ac = [0], [], []
    for j in range(n):
        nz = np.nonzero(A[:, j])[0]
        mx = np.abs(A[nz, j]).max() if len(nz) else 1.0
        ri += list(nz)
        ac += list(A[nz, j] / mx)
        cp.append(len(ri))
    g = go.Graph(m, n, np.array(rp), np.array(ci), np.array(ar), np.array(cp), np.array(ri), np.array(ac))
    Xc = rng.standard_normal((n, go.DIN))
    Xr = rng.standard_normal((m, go.DIN))
    tx = rng.standard_normal(n)
    ty = rng.standard_normal(m)
    return g, Xc, Xr, tx, ty


def check_finite_differences(go, flat, idx, loss_at, gflat, names):
    """Central differences at every index in idx, except where the point
    sits on a ReLU kink (the one-sided differences disagree by more than
    their truncation and rounding allow): there the derivative is not
    defined.  At most 5 % of the points may be kinks.  The rounding of the
    loss itself (about 10 ulps of |loss|, divided by h) enters the bound."""
    h = 1e-5
    f0 = loss_at(flat)
    rnd = 10 * np.finfo(np.float64).eps * abs(f0) / h
    kinks = 0
    for k in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[k] += h
        fm[k] -= h
        lp_, lm_ = loss_at(fp), loss_at(fm)
        fwd, bwd = (lp_ - f0) / h, (f0 - lm_) / h
        if abs(fwd - bwd) > 1e-3 * (abs(fwd) + abs(bwd)) + 4 * rnd + 1e-7:
            kinks += 1
            continue
        fd = (lp_ - lm_) / (2 * h)
        assert abs(fd - gflat[k]) <= 1e-6 + 1e-5 * abs(fd) + rnd, (names[k], fd, gflat[k])
    assert kinks <= 0.05 * len(idx), kinks


def param_names(go):
    return [name for name, shp in go.param_shapes() for _ in range(int(np.prod(shp)))]


def test_gradients_match_finite_differences():
    from oracle import gnn_oracle as go
    g, Xc, Xr, tx, ty = tiny_graph(go)
    rng = np.random.default_rng(1)
    flat = rng.standard_normal(go.n_params()) * 0.3
    p = go.unflatten(flat)
    loss, gr, _ = go.loss_and_grad(p, g, Xc, Xr, tx, ty, use_bf16=False)
    idx = rng.choice(len(flat), 120, replace=False)
    check_finite_differences(go, flat, idx,
                             lambda f: go.loss_and_grad(go.unflatten(f), g, Xc, Xr, tx, ty, use_bf16=False)[0],
                             go.flatten(gr), param_names(go))


def test_primal_weight_head_matches_finite_differences():
    """The same check with the primal-weight head in the loss, over a graph
    cut into three LPs (one of them without columns)."""
    from oracle import gnn_oracle as go
    g, Xc, Xr, tx, ty = tiny_graph(go, seed=4)
    seg = np.array([0, 4, 4, g.n])
    tw = np.array([0.7, -0.3, 1.9])
    rng = np.random.default_rng(5)
    flat = rng.standard_normal(go.n_params()) * 0.3
    loss, gr, c = go.loss_and_grad(go.unflatten(flat), g, Xc, Xr, tx, ty, use_bf16=False, seg=seg, tw=tw)
    # the pooled embedding of the empty LP is zero, so its prediction is the bias
    assert c["what"][1] == go.unflatten(flat)["bom"][0]
    names = param_names(go)
    head = [k for k, nm in enumerate(names) if nm in ("wom", "bom")]
    idx = list(rng.choice(len(flat), 80, replace=False)) + head
    check_finite_differences(
        go, flat, idx,
        lambda f: go.loss_and_grad(go.unflatten(f), g, Xc, Xr, tx, ty, use_bf16=False, seg=seg, tw=tw)[0],
        go.flatten(gr), names)


def test_bf16_rounding():
    from oracle import gnn_oracle as go
    # exactly representable values survive; ties round to even
    assert go.bf16(np.array([1.0, -2.5, 0.0]))[1] == -2.5
    one = np.float32(1.0).view(np.uint32)
    tie_down = np.array([one + 0x8000], np.uint32).view(np.float32)  # halfway, even below
    tie_up = np.array([one + 0x18000], np.uint32).view(np.float32)   # halfway, odd below
    assert go.bf16(tie_down)[0] == 1.0
    assert go.bf16(tie_up)[0] == np.array([one + 0x20000], np.uint32).view(np.float32)[0]
    x = np.random.default_rng(0).standard_normal(1000)
    assert np.max(np.abs(go.bf16(x) - x) / np.abs(x)) <= 2.0 ** -8


def test_adam_closed_form():
    from oracle import gnn_oracle as go
    p, m1, m2 = go.adam(np.array([1.0]), np.array([0.5]), np.zeros(1), np.zeros(1), 1, lr=0.1)
    # first step: m_hat = g, v_hat = g^2 -> update = lr * sign(g) (up to eps)
    assert p[0] == pytest.approx(0.9, abs=1e-7)
