/*
 * pramaan_gnn.h -- the learned warm start: a bipartite row/column
 * message-passing network (2 layers, hidden 64) that predicts a primal and
 * dual starting point and the initial primal weight for an LP.  Dense 64x64 layers run on MFMA in
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
 *   We_c[16][64] be_c[64] We_r[16][64] be_r[64]
 *   for l = 1, 2: Wrr[64][64] Wrz[64][64] br[64] Wcc[64][64] Wcz[64][64] bc[64]
 *   wx[64] bx[1] wy[64] by[1] wom[64] bom[1]
 * (row-major; W[k][c] multiplies input feature k into output feature c).
 * The primal-weight head predicts, for each LP k of a (block-diagonal)
 * graph, log(final primal weight / default primal weight) as
 * mean_{columns j of k} Hc2[j] . wom + bom.
 */
#ifndef PRAMAAN_GNN_H
#define PRAMAAN_GNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_GNN_HIDDEN 64
#define PM_GNN_FEATURES 16
#define PM_GNN_ROLE_BITS 9

struct pm_lp;

/* Host-side graph and features of one LP (no device work):
 *   cp[n+1], ri[nnz], ac[nnz]  A^T in CSR (rows ascending per column), each
 *                              column scaled to unit max-norm;
 *   ar[nnz]                    A's values (CSR order of lp) with each row
 *                              scaled to unit max-norm;
 *   xc[n][16] column features  c/cs, l/s, u/s (0 when infinite), l finite,
 *                              u finite, nnz/max nnz, period position, then
 *                              the 9 role bits;
 *   xr[m][16] row features     l/s, u/s, l finite, u finite, equality,
 *                              nnz/max nnz, period position, role bits;
 *   scales[2] = {s, cs}        s = max(1, largest finite |bound|),
 *                              cs = max(1e-12, max |c|).
 * The period position is (index / period) / (count / period - 1) when a
 * period length is given (> 0), else 0.  The role bits say which node of
 * its period a node is: bit b (b < 9) of (index mod period) as -1 or +1,
 * all 0 without a period.  Targets of the network are x / s and y / cs. */
int pm_gnn_lp_features(const struct pm_lp* lp, int32_t period_cols,
                       int32_t period_rows, int32_t* cp, int32_t* ri,
                       float* ar, float* ac, float* xr, float* xc,
                       double* scales);

typedef struct pm_gnn pm_gnn;

/* Number of FP32 parameters. */
int32_t pm_gnn_num_params(void);
/* Creates a network with deterministic scaled-uniform initial weights. */
int pm_gnn_create(int32_t device, uint64_t seed, pm_gnn** out);
int pm_gnn_destroy(pm_gnn* g);
/* Page-locked host memory (no counterpart in the reference; for callers
 * that build graphs for pm_gnn_set_graph and want their copies to run at
 * full link speed).  pm_host_free(NULL) does nothing. */
int pm_host_alloc(size_t bytes, void** out);
int pm_host_free(void* p);
int pm_gnn_get_params(pm_gnn* g, float* out);
int pm_gnn_set_params(pm_gnn* g, const float* in);  /* also resets Adam */

/* The graph of one batch of LPs (block-diagonal if several): Ar = A with
 * rows scaled to unit max-norm, in CSR (rp[m+1], ci, ar); Ac = A^T with
 * rows (= columns of A) scaled to unit max-norm, in CSR (cp[n+1], ri, ac);
 * row features xr[m][PM_GNN_FEATURES], column features
 * xc[n][PM_GNN_FEATURES].  Copied to the device, where the indices are
 * checked: in range, strictly increasing within each row of Ar and each
 * row of Ac, and Ac holding exactly the entries of Ar.  A graph failing
 * them is refused (PM_ERR_INVALID_ARG) and the network's graph cleared.
 * Arrays in page-locked memory (pm_host_alloc) copy at full link speed. */
int pm_gnn_set_graph(pm_gnn* g, int32_t m, int32_t n, int32_t nnz,
                     const int32_t* rp, const int32_t* ci, const float* ar,
                     const int32_t* cp, const int32_t* ri, const float* ac,
                     const float* xr, const float* xc);
/* Predictions for the current graph: xhat[n], yhat[m]. */
int pm_gnn_forward(pm_gnn* g, float* xhat, float* yhat);
/* The LPs of the current graph for the primal-weight head: LP k owns the
 * columns [col_off[k], col_off[k+1]) (count + 1 offsets from 0 to n).
 * pm_gnn_set_graph resets this to one LP holding every column. */
int pm_gnn_set_segments(pm_gnn* g, int32_t count, const int32_t* col_off);
/* Training targets of the primal-weight head, one per LP (log of the final
 * over the default primal weight); with them the loss gains 0.01 x the
 * mean squared error of the head.  NULL removes them.  Cleared by
 * pm_gnn_set_graph and pm_gnn_set_segments. */
int pm_gnn_set_omega_targets(pm_gnn* g, const float* targets);
/* Forward pass and the head's predictions, one per LP. */
int pm_gnn_forward_omega(pm_gnn* g, float* omega_hat);
/* Whether pm_batch_learned_warm_start applies the head (default 0: on the
 * rolling-horizon windows the predicted weight made the learned solve
 * slower, 3 % on the measured sequence and 10 % on a held-out one; the
 * head's target, the weight a cold solve ends with, is not the best
 * weight to start a warm solve with). */
int pm_gnn_set_omega_head(pm_gnn* g, int32_t enabled);
/* Loss (mean squared error over columns plus over rows, plus the head's
 * when targets are set) and its gradient w.r.t. the flat parameters,
 * without updating them. */
int pm_gnn_gradient(pm_gnn* g, const float* tx, const float* ty, float* loss,
                    float* grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the current graph. */
int pm_gnn_train_step(pm_gnn* g, const float* tx, const float* ty, float lr,
                      float* loss);

/* Learned warm start of a resident batch (pramaan_hip.h): the graph and
 * features of every LP of the batch are built on the device from the
 * batch's own arrays, exactly as pm_gnn_lp_features builds them on the host
 * (period lengths as there), one forward pass runs over the block-diagonal
 * graph of all LPs, and x0 = clamp(xhat * s, l, u), y0 = yhat * cs become
 * the warm starts of every LP for the following pm_batch_run, which judges
 * each against the cold start; an LP whose warm start is kept starts from
 * the default primal weight, or, with pm_gnn_set_omega_head(g, 1), the
 * default times exp(the head's prediction), the factor kept within
 * [1/100, 100] (pm_batch_set_primal_weight still fixes the weight).  The network's current graph is replaced
 * (call pm_gnn_set_graph again before training).  *seconds (may be NULL)
 * receives the device time of the whole step. */
struct pm_batch;
int pm_batch_learned_warm_start(struct pm_batch* b, pm_gnn* g,
                                int32_t period_cols, int32_t period_rows,
                                double* seconds);
/* The warm start LP i of the batch will use (original space, y in the
 * reference sign); fails if it has none. */
int pm_batch_get_warm_start(struct pm_batch* b, int32_t i, double* x0,
                            double* y0);

#ifdef __cplusplus
}
#endif
#endif
