// gnn_internal.hpp -- the warm-start network's state (pm_gnn), shared by
// gnn.hip and the learned warm start of a resident batch (learned.hip).
#pragma once

#include <vector>

#include "common.hpp"
#include "pramaan_gnn.h"

struct pm_gnn {
    int device = 0;
    hipStream_t s = nullptr;
    int step = 0;
    int m = 0, n = 0, nnz = 0;
    bool grad_ready = false;  // graph set by pm_gnn_set_graph (backward buffers and A^T values present)
    pm::DevBuf<float> p, grad, m1, m2;
    pm::DevBuf<short> wt, wn;  // 8 dense weights x 4096 bf16 each, two layouts
    pm::DevBuf<int> rp, ci, cp, ri;
    pm::DevBuf<float> ar, ac, ar_t, ac_t, xr, xc;
    pm::DevBuf<int> flag;  // graph checks of pm_gnn_set_graph
    // activations Hc0..2, Hr0..2, Zr1..2, Zc1..2 stored as bfloat16 bits
    // (every consumer reads them in bfloat16: the MFMA layers, the message
    // passing, the heads); gradients and everything else in FP32
    pm::DevBuf<short> Hc[3], Hr[3], Zr[3], Zc[3];
    pm::DevBuf<float> dHc, dHr, dHc2, dZ, dT, xhat, yhat, dx, dy, tx, ty, part;
    // primal-weight head: the graph's LPs own the columns [seg[k], seg[k+1]);
    // pool[k][64] = mean of their final column embeddings, what[k] the
    // prediction, tw[k] its training target (used when has_tw), dw[k] the
    // loss derivative.
    int nseg = 1;
    bool has_tw = false;
    bool omega_head = false;  // pm_batch_learned_warm_start applies what (pm_gnn_set_omega_head)
    pm::DevBuf<int> seg;
    pm::DevBuf<float> pool, what, tw, dw;
    std::vector<float> host_tmp;
};

namespace pm {
// Device buffers for a graph of m rows, n columns and nnz edges (the
// gradient buffers too when training); the graph arrays are not touched.
void gnn_reserve(pm_gnn* g, int m, int n, int nnz, bool training);
// Forward pass over the graph in g->rp ... g->xc into g->xhat, g->yhat and
// the primal-weight predictions g->what (one per segment).
void gnn_forward(pm_gnn* g);
// Segments of the primal-weight head: count LPs whose columns start at
// col_off[k] (host array of count + 1, from 0 to n).
void gnn_set_segments(pm_gnn* g, int count, const int* col_off);
}  // namespace pm
