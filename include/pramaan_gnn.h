/*
 * pramaan_gnn.h -- the learned warm start: a bipartite row/column
 * message-passing network (2 layers, hidden 64) that predicts a primal and
 * dual starting point for an LP.  Dense 64x64 layers run on MFMA in
 * bfloat16 with FP32 accumulation; message passing (sparse aggregation over
 * the constraint matrix), heads, loss and Adam run in FP32.  Forward,
 * backward and Adam are hand-written HIP kernels.
 *
 * No reference counterpart: the reference warm-starts only its simplex
 * from a stored basis (include/pramaan/dual_simplex.hpp:42-126).  The
 * network definition (features, layers, loss) is documented in
 * oracle/gnn_oracle.py, which the tests check the kernels against.
 *
 * Parameters are one flat FP32 vector in this order:
 *   We_c[8][64] be_c[64] We_r[8][64] be_r[64]
 *   for l = 1, 2: Wrr[64][64] Wrz[64][64] br[64] Wcc[64][64] Wcz[64][64] bc[64]
 *   wx[64] bx[1] wy[64] by[1]
 * (row-major; W[k][c] multiplies input feature k into output feature c).
 */
#ifndef PRAMAAN_GNN_H
#define PRAMAAN_GNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_GNN_HIDDEN 64
#define PM_GNN_FEATURES 8

typedef struct pm_gnn pm_gnn;

/* Number of FP32 parameters. */
int32_t pm_gnn_num_params(void);
/* Creates a network with deterministic scaled-uniform initial weights. */
int pm_gnn_create(int32_t device, uint64_t seed, pm_gnn** out);
int pm_gnn_destroy(pm_gnn* g);
int pm_gnn_get_params(pm_gnn* g, float* out);
int pm_gnn_set_params(pm_gnn* g, const float* in);  /* also resets Adam */

/* The graph of one batch of LPs (block-diagonal if several): Ar = A with
 * rows scaled to unit max-norm, in CSR (rp[m+1], ci, ar); Ac = A^T with
 * rows (= columns of A) scaled to unit max-norm, in CSR (cp[n+1], ri, ac);
 * row features xr[m][8], column features xc[n][8].  Copied to the device. */
int pm_gnn_set_graph(pm_gnn* g, int32_t m, int32_t n, int32_t nnz,
                     const int32_t* rp, const int32_t* ci, const float* ar,
                     const int32_t* cp, const int32_t* ri, const float* ac,
                     const float* xr, const float* xc);
/* Predictions for the current graph: xhat[n], yhat[m]. */
int pm_gnn_forward(pm_gnn* g, float* xhat, float* yhat);
/* Loss (mean squared error over columns plus over rows) and its gradient
 * w.r.t. the flat parameters, without updating them. */
int pm_gnn_gradient(pm_gnn* g, const float* tx, const float* ty, float* loss,
                    float* grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the current graph. */
int pm_gnn_train_step(pm_gnn* g, const float* tx, const float* ty, float lr,
                      float* loss);

#ifdef __cplusplus
}
#endif
#endif
