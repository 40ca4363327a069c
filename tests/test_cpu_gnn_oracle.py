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


def test_gradients_match_finite_differences():
    from oracle import gnn_oracle as go
    g, Xc, Xr, tx, ty = tiny_graph(go)
    rng = np.random.default_rng(1)
    flat = rng.standard_normal(go.n_params()) * 0.3
    p = go.unflatten(flat)
    loss, gr, _ = go.loss_and_grad(p, g, Xc, Xr, tx, ty, use_bf16=False)
    gflat = go.flatten(gr)
    idx = rng.choice(len(flat), 120, replace=False)
    h = 1e-6
    for k in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[k] += h
        fm[k] -= h
        lp_ = go.loss_and_grad(go.unflatten(fp), g, Xc, Xr, tx, ty, use_bf16=False)[0]
        lm_ = go.loss_and_grad(go.unflatten(fm), g, Xc, Xr, tx, ty, use_bf16=False)[0]
        fd = (lp_ - lm_) / (2 * h)
        assert abs(fd - gflat[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, gflat[k])


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
