/*
 * pdhg_oracle.h -- CPU oracle for the GPU PDHG path.  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline.
 * The product path never calls it.
 *
 * It restates, in plain C, the reference's algorithm for this path
 * (paths relative to the reference repo):
 *   or_spmv_f32          csr_spmv_kernel        src/gpu/pdhg_kernels.cuh:24-37
 *   or_spmv_t_f32        csr_spmv_t_kernel      src/gpu/pdhg_kernels.cuh:42-54
 *                        (row-order scatter; the reference's atomic order is
 *                        unspecified, this is the sequential order)
 *   or_dual_update       dual_update_kernel     src/gpu/pdhg_kernels.cuh:72-89
 *   or_primal_update     primal_update_kernel   src/gpu/pdhg_kernels.cuh:95-109
 *   or_step1_iteration   pdhg_primal/dual_update_kernel + launch_pdhg_iteration
 *                                               src/gpu/pdhg.cu:42-66,77-146
 *   or_fp64_residuals    ComputeFp64Residuals   src/gpu/pdhg_solver.cu:107-161
 *   or_pdhg_solve        PdhgSolver::solve      src/gpu/pdhg_solver.cu:169-411
 *   or_transpose         CSRMatrix::transpose   src/linalg/sparse_matrix.cpp:13-57
 *   or_ruiz_scale        ruiz_scale             src/presolve/scaling.cpp:29-150
 * and adds the FP64 relative KKT error (PDLP definition) the engine's
 * PM_TERM_RELATIVE_KKT test uses:
 *   or_relative_kkt
 * Built with -ffp-contract=off so every a*b+c is two rounded operations,
 * matching the reference's source expressions.
 */
#ifndef PDHG_ORACLE_H
#define PDHG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_lp {
    int32_t m, n, nnz;
    int32_t sense; /* 0 min, 1 max */
    double offset;
    const double* c;
    const int32_t* row_ptr;
    const int32_t* col_idx;
    const double* values;
    const double* row_lower;
    const double* row_upper;
    const double* var_lower;
    const double* var_upper;
} or_lp;

typedef struct or_result {
    int32_t status;     /* 0 optimal, 3 iteration limit, 4 numerical failure */
    int32_t iterations;
    int32_t restarts;
    double objective;
    double primal_residual;
    double dual_residual;
} or_result;

void or_spmv_f32(int32_t m, const int32_t* rp, const int32_t* ci,
                 const float* v, const float* x, float* y);
void or_spmv_t_f32(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci,
                   const float* v, const float* x, float* y);
void or_dual_update(int32_t m, const float* ax_bar, const float* sigma,
                    const float* l_r, const float* u_r, float* y);
void or_primal_update(int32_t n, const float* aty, const float* tau,
                      const float* c, const float* l_x, const float* u_x,
                      float* x, float* x_bar);
void or_step1_iteration(float* x, float* y, const float* gradient,
                        const float* ax_bar, const float* b, int32_t n,
                        int32_t m, float tau, float sigma);
void or_fp64_residuals(const or_lp* lp, const float* x, const float* y,
                       const float* c32, const float* v32, double tol,
                       double* pv, double* dv);
/* Full reference solve.  x_out[n], y_out[m] (FP32 iterates) may be NULL. */
int or_pdhg_solve(const or_lp* lp, double tol, int32_t max_iterations,
                  int32_t check_frequency, or_result* res, float* x_out,
                  float* y_out);
void or_transpose(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci,
                  const double* v, int32_t* t_rp, int32_t* t_ci, double* t_v);
void or_ruiz_scale(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci,
                   double* v, double* c, double* row_lower, double* row_upper,
                   double* var_lower, double* var_upper, int32_t passes,
                   double* row_scale, double* col_scale);
/* Relative KKT of (x, y) for the LP in minimisation form (c negated for a
 * maximisation), y in the reference sign (reduced cost c + A^T y).
 * out[0..6] = rel_kkt, rel_primal, rel_dual, rel_gap, primal_obj, dual_obj,
 *             ||A x - proj(A x)||_inf.  */
void or_relative_kkt(const or_lp* lp, const double* x, const double* y,
                     double* out);

#ifdef __cplusplus
}
#endif
#endif
