"""CPU oracle of the learned warm-start network (numpy, FP64).  TEST
INFRASTRUCTURE: used by tests only, to check the HIP forward, backward and
Adam kernels.  No reference counterpart exists (the reference has no
learned component); the network is specified here and in
pramaan_amd/csrc/gnn.hip, and this file pins it by finite differences
(tests/test_cpu_gnn_oracle.py) before it is used to check the GPU.

Network (bipartite row/column message passing, hidden H = 64):
  Hc0 = relu(Xc We_c + be_c)              Hr0 = relu(Xr We_r + be_r)
  for l in 1, 2:
    Zr = Ar Hc(l-1);  Hr(l) = relu(Hr(l-1) Wrr_l + Zr Wrz_l + br_l)
    Zc = Ac Hr(l);    Hc(l) = relu(Hc(l-1) Wcc_l + Zc Wcz_l + bc_l)
  xhat = Hc2 wx + bx;  yhat = Hr2 wy + by
  primal-weight head, per LP g of a block-diagonal graph whose columns
  [seg[g], seg[g+1]) belong to it:  what_g = mean_{j in g} Hc2[j] . wom + bom
  loss = mean_j (xhat_j - tx_j)^2 + mean_i (yhat_i - ty_i)^2
         (+ LW * mean_g (what_g - tw_g)^2 when targets tw are given, LW = 0.01:
          Adam scales each parameter's step, so the head learns at full speed
          while its term moves the shared layers little)
Ar is A with rows scaled by 1/max|row| (m x n); Ac is A^T with rows scaled
by 1/max|column| (n x m).  With bf16=True every activation (Hc, Hr, Zr, Zc)
is stored rounded to bfloat16 (round to nearest even) where it is produced,
as the HIP kernels store them, and every 64x64 product rounds its two inputs
to bfloat16 and accumulates in FP64, as the MFMA kernels do with FP32
accumulation.  Gradients stay unrounded except as inputs of those products
(the rounding passes gradients through unchanged).
"""
from __future__ import annotations

import numpy as np

H = 64
LW = 0.01  # weight of the primal-weight head's term in the loss
DIN = 16  # node features: 7 from the LP and 9 role bits (include/pramaan_gnn.h)
LAYER_NAMES = ["Wrr", "Wrz", "br", "Wcc", "Wcz", "bc"]


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float64/float32 values to bfloat16 (nearest even), return FP64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def param_shapes():
    shapes = [("We_c", (DIN, H)), ("be_c", (H,)), ("We_r", (DIN, H)), ("be_r", (H,))]
    for l in (1, 2):
        shapes += [(f"Wrr{l}", (H, H)), (f"Wrz{l}", (H, H)), (f"br{l}", (H,)),
                   (f"Wcc{l}", (H, H)), (f"Wcz{l}", (H, H)), (f"bc{l}", (H,))]
    shapes += [("wx", (H,)), ("bx", (1,)), ("wy", (H,)), ("by", (1,)), ("wom", (H,)), ("bom", (1,))]
    return shapes


def unflatten(flat):
    out, o = {}, 0
    for name, shp in param_shapes():
        k = int(np.prod(shp))
        out[name] = np.asarray(flat[o:o + k], dtype=np.float64).reshape(shp)
        o += k
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in param_shapes()])


def n_params():
    return sum(int(np.prod(s)) for _, s in param_shapes())


class Graph:
    """CSR matrices of the aggregation: Ar (m x n) and Ac (n x m)."""

    def __init__(self, m, n, rp, ci, ar, cp, ri, ac):
        self.m, self.n = m, n
        self.rp, self.ci, self.ar = rp, ci, np.asarray(ar, np.float64)
        self.cp, self.ri, self.ac = cp, ri, np.asarray(ac, np.float64)

    def row_agg(self, Hc):  # Ar @ Hc
        out = np.zeros((self.m, Hc.shape[1]))
        for i in range(self.m):
            s, e = self.rp[i], self.rp[i + 1]
            out[i] = self.ar[s:e] @ Hc[self.ci[s:e]]
        return out

    def col_agg(self, Hr):  # Ac @ Hr
        out = np.zeros((self.n, Hr.shape[1]))
        for j in range(self.n):
            s, e = self.cp[j], self.cp[j + 1]
            out[j] = self.ac[s:e] @ Hr[self.ri[s:e]]
        return out

    def row_agg_T(self, D):  # Ar^T @ D (n x H)
        out = np.zeros((self.n, D.shape[1]))
        for i in range(self.m):
            s, e = self.rp[i], self.rp[i + 1]
            np.add.at(out, self.ci[s:e], np.outer(self.ar[s:e], D[i]))
        return out

    def col_agg_T(self, D):  # Ac^T @ D (m x H)
        out = np.zeros((self.m, D.shape[1]))
        for j in range(self.n):
            s, e = self.cp[j], self.cp[j + 1]
            np.add.at(out, self.ri[s:e], np.outer(self.ac[s:e], D[j]))
        return out


def _mm(a, b, use_bf16):
    return (bf16(a) @ bf16(b)) if use_bf16 else (a @ b)


def forward(p, g: Graph, Xc, Xr, use_bf16=True, seg=None):
    relu = lambda t: np.maximum(t, 0.0)
    st = bf16 if use_bf16 else (lambda t: t)  # activation storage
    c = {}
    c["Hc0"] = st(relu(Xc @ p["We_c"] + p["be_c"]))
    c["Hr0"] = st(relu(Xr @ p["We_r"] + p["be_r"]))
    Hc, Hr = c["Hc0"], c["Hr0"]
    for l in (1, 2):
        Zr = st(g.row_agg(Hc))
        Hr = st(relu(_mm(Hr, p[f"Wrr{l}"], use_bf16) + _mm(Zr, p[f"Wrz{l}"], use_bf16) + p[f"br{l}"]))
        Zc = st(g.col_agg(Hr))
        Hc = st(relu(_mm(Hc, p[f"Wcc{l}"], use_bf16) + _mm(Zc, p[f"Wcz{l}"], use_bf16) + p[f"bc{l}"]))
        c[f"Zr{l}"], c[f"Hr{l}"], c[f"Zc{l}"], c[f"Hc{l}"] = Zr, Hr, Zc, Hc
    c["xhat"] = Hc @ p["wx"] + p["bx"][0]
    c["yhat"] = Hr @ p["wy"] + p["by"][0]
    seg = np.array([0, g.n]) if seg is None else np.asarray(seg)
    c["pool"] = np.stack([Hc[seg[k]:seg[k + 1]].mean(axis=0) if seg[k + 1] > seg[k] else np.zeros(Hc.shape[1])
                          for k in range(len(seg) - 1)])
    c["what"] = c["pool"] @ p["wom"] + p["bom"][0]
    return c


def loss_and_grad(p, g: Graph, Xc, Xr, tx, ty, use_bf16=True, seg=None, tw=None):
    c = forward(p, g, Xc, Xr, use_bf16, seg)
    n, m = g.n, g.m
    ex, ey = c["xhat"] - tx, c["yhat"] - ty
    loss = float(np.mean(ex ** 2) + np.mean(ey ** 2))
    dx, dy = 2.0 * ex / n, 2.0 * ey / m
    gr = {k: np.zeros_like(v) for k, v in p.items()}
    gr["wx"] = c["Hc2"].T @ dx
    gr["bx"] = np.array([dx.sum()])
    gr["wy"] = c["Hr2"].T @ dy
    gr["by"] = np.array([dy.sum()])
    dHc = np.outer(dx, p["wx"])
    dHr = np.outer(dy, p["wy"])
    if tw is not None:
        seg = np.array([0, n]) if seg is None else np.asarray(seg)
        ew = c["what"] - np.asarray(tw, np.float64)
        loss += LW * float(np.mean(ew ** 2))
        dw = 2.0 * LW * ew / len(ew)
        gr["wom"] = c["pool"].T @ dw
        gr["bom"] = np.array([dw.sum()])
        for k in range(len(seg) - 1):
            cnt = seg[k + 1] - seg[k]
            if cnt > 0:
                dHc[seg[k]:seg[k + 1]] += dw[k] * p["wom"] / cnt
    for l in (2, 1):
        Hc_prev = c[f"Hc{l - 1}"]
        Hr_prev = c[f"Hr{l - 1}"]
        # column update of layer l
        dP = dHc * (c[f"Hc{l}"] > 0)
        gr[f"Wcc{l}"] = _mm(Hc_prev.T, dP, use_bf16)
        gr[f"Wcz{l}"] = _mm(c[f"Zc{l}"].T, dP, use_bf16)
        gr[f"bc{l}"] = dP.sum(axis=0)
        dHc_prev = _mm(dP, p[f"Wcc{l}"].T, use_bf16)
        dZc = _mm(dP, p[f"Wcz{l}"].T, use_bf16)
        dHr = dHr + g.col_agg_T(dZc)
        # row update of layer l
        dP = dHr * (c[f"Hr{l}"] > 0)
        gr[f"Wrr{l}"] = _mm(Hr_prev.T, dP, use_bf16)
        gr[f"Wrz{l}"] = _mm(c[f"Zr{l}"].T, dP, use_bf16)
        gr[f"br{l}"] = dP.sum(axis=0)
        dHr = _mm(dP, p[f"Wrr{l}"].T, use_bf16)
        dZr = _mm(dP, p[f"Wrz{l}"].T, use_bf16)
        dHc = dHc_prev + g.row_agg_T(dZr)
    dP = dHc * (c["Hc0"] > 0)
    gr["We_c"] = Xc.T @ dP
    gr["be_c"] = dP.sum(axis=0)
    dP = dHr * (c["Hr0"] > 0)
    gr["We_r"] = Xr.T @ dP
    gr["be_r"] = dP.sum(axis=0)
    return loss, gr, c


def adam(flat, grad, m1, m2, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, 2015) on flat FP64 arrays; returns new
    (params, m1, m2).  step counts from 1."""
    m1 = b1 * m1 + (1 - b1) * grad
    m2 = b2 * m2 + (1 - b2) * grad * grad
    mh = m1 / (1 - b1 ** step)
    vh = m2 / (1 - b2 ** step)
    return flat - lr * mh / (np.sqrt(vh) + eps), m1, m2
