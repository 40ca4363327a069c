/*
 * pdhg_oracle.c -- CPU oracle for the GPU PDHG path (test infrastructure
 * only; see pdhg_oracle.h for the rules and the reference citations).
 *
 * Every function restates the reference step by step, in the same
 * floating-point order, so that a GPU kernel run in its sequential variant
 * can be compared bit for bit.  Paths are relative to the reference repo.
 */
#include "pdhg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_INF 1.0e30          /* kInfinity, include/pramaan/ir.hpp:47      */
#define OR_FLT_INF 1.0e30f     /* kFltInf, src/gpu/pdhg_solver.cu:200       */

static float fmin_f(float a, float b) { return fminf(a, b); }
static float fmax_f(float a, float b) { return fmaxf(a, b); }

/* csr_spmv_kernel, pdhg_kernels.cuh:24-37: y_i = sum_k v_k x_{col_k},
 * accumulated in row order starting from 0.0f. */
void or_spmv_f32(int32_t m, const int32_t* rp, const int32_t* ci,
                 const float* v, const float* x, float* y) {
    for (int32_t i = 0; i < m; ++i) {
        float sum = 0.0f;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) sum += v[k] * x[ci[k]];
        y[i] = sum;
    }
}

/* csr_spmv_t_kernel, pdhg_kernels.cuh:42-54: y[col_k] += v_k x_i over a
 * zeroed y (pdhg_solver.cu:332).  Rows in increasing order, so each y_j
 * sums its terms in increasing row index. */
void or_spmv_t_f32(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci,
                   const float* v, const float* x, float* y) {
    for (int32_t j = 0; j < n; ++j) y[j] = 0.0f;
    for (int32_t i = 0; i < m; ++i) {
        const float xi = x[i];
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) y[ci[k]] += v[k] * xi;
    }
}

/* dual_update_kernel, pdhg_kernels.cuh:72-89. */
void or_dual_update(int32_t m, const float* ax_bar, const float* sigma,
                    const float* l_r, const float* u_r, float* y) {
    for (int32_t i = 0; i < m; ++i) {
        const float sig = sigma[i];
        if (sig <= 0.0f) continue;
        const float w = y[i] / sig + ax_bar[i];
        const float w_proj = fmin_f(fmax_f(w, l_r[i]), u_r[i]);
        y[i] = sig * (w - w_proj);
    }
}

/* primal_update_kernel, pdhg_kernels.cuh:95-109. */
void or_primal_update(int32_t n, const float* aty, const float* tau,
                      const float* c, const float* l_x, const float* u_x,
                      float* x, float* x_bar) {
    for (int32_t j = 0; j < n; ++j) {
        const float x_prev = x[j];
        const float x_new =
            fmin_f(fmax_f(x_prev - tau[j] * (c[j] + aty[j]), l_x[j]), u_x[j]);
        x_bar[j] = 2.0f * x_new - x_prev;
        x[j] = x_new;
    }
}

/* launch_pdhg_iteration with its two Step-1 kernels, pdhg.cu:42-66,77-146:
 * x <- max(0, x - tau g) and y <- y + sigma (Ax_bar - b), in place. */
void or_step1_iteration(float* x, float* y, const float* gradient,
                        const float* ax_bar, const float* b, int32_t n,
                        int32_t m, float tau, float sigma) {
    if (n <= 0 && m <= 0) return;
    for (int32_t j = 0; j < n; ++j) x[j] = fmax_f(0.0f, x[j] - tau * gradient[j]);
    for (int32_t i = 0; i < m; ++i) y[i] = y[i] + sigma * (ax_bar[i] - b[i]);
}

/* ComputeFp64Residuals, pdhg_solver.cu:107-161.  A x uses the model's FP64
 * values (the reference calls model.A.multiply); A^T y uses the FP32
 * values v32 the device holds; c32 is the FP32 minimisation-form cost. */
void or_fp64_residuals(const or_lp* lp, const float* x, const float* y,
                       const float* c32, const float* v32, double tol,
                       double* out_pv, double* out_dv) {
    const int32_t m = lp->m, n = lp->n;
    double* xd = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* aty = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int32_t j = 0; j < n; ++j) xd[j] = (double)x[j];
    double pv = 0.0;
    for (int32_t i = 0; i < m; ++i) {
        double act = 0.0;
        for (int32_t k = lp->row_ptr[i]; k < lp->row_ptr[i + 1]; ++k)
            act += lp->values[k] * xd[lp->col_idx[k]];
        if (act < lp->row_lower[i] && lp->row_lower[i] - act > pv) pv = lp->row_lower[i] - act;
        if (act > lp->row_upper[i] && act - lp->row_upper[i] > pv) pv = act - lp->row_upper[i];
    }
    for (int32_t j = 0; j < n; ++j) {
        if (xd[j] < lp->var_lower[j] && lp->var_lower[j] - xd[j] > pv) pv = lp->var_lower[j] - xd[j];
        if (xd[j] > lp->var_upper[j] && xd[j] - lp->var_upper[j] > pv) pv = xd[j] - lp->var_upper[j];
    }
    for (int32_t i = 0; i < m; ++i) {
        const double yi = (double)y[i];
        for (int32_t k = lp->row_ptr[i]; k < lp->row_ptr[i + 1]; ++k)
            aty[lp->col_idx[k]] += (double)v32[k] * yi;
    }
    double c_norm = 1.0;
    for (int32_t j = 0; j < n; ++j) {
        const double a = (double)fabsf(c32[j]);
        if (a > c_norm) c_norm = a;
    }
    double dv = 0.0;
    for (int32_t j = 0; j < n; ++j) {
        const double rc = (double)c32[j] + aty[j];
        const int at_lb = xd[j] <= lp->var_lower[j] + tol;
        const int at_ub = xd[j] >= lp->var_upper[j] - tol;
        double viol = 0.0;
        if (!at_lb && !at_ub) viol = fabs(rc);
        else if (at_lb && !at_ub) viol = (-rc > 0.0) ? -rc : 0.0;
        else if (at_ub && !at_lb) viol = (rc > 0.0) ? rc : 0.0;
        if (viol > dv) dv = viol; /* std::max(dv, viol): a NaN never wins */
    }
    dv /= c_norm;
    *out_pv = pv;
    *out_dv = dv;
    free(xd);
    free(aty);
}

/* PdhgSolver::solve, pdhg_solver.cu:169-411 (restated step by step). */
int or_pdhg_solve(const or_lp* lp, double tol, int32_t max_iterations,
                  int32_t check_frequency, or_result* res, float* x_out,
                  float* y_out) {
    const int32_t m = lp->m, n = lp->n, nnz = lp->nnz;
    memset(res, 0, sizeof(*res));
    res->status = 3;
    res->primal_residual = INFINITY;
    res->dual_residual = INFINITY;
    if (n == 0) { /* 178-184 */
        res->status = 0;
        res->objective = lp->offset;
        return 0;
    }
    for (int32_t j = 0; j < n; ++j) /* 187-192 */
        if (!isfinite(lp->c[j])) { res->status = 4; return 0; }

    const float obj_sign = lp->sense == 1 ? -1.0f : 1.0f; /* 196 */
    size_t nz = (size_t)(nnz > 0 ? nnz : 1), ms = (size_t)(m > 0 ? m : 1);
    float* c = (float*)malloc(sizeof(float) * (size_t)n);
    float* l_r = (float*)malloc(sizeof(float) * ms);
    float* u_r = (float*)malloc(sizeof(float) * ms);
    float* l_x = (float*)malloc(sizeof(float) * (size_t)n);
    float* u_x = (float*)malloc(sizeof(float) * (size_t)n);
    float* v = (float*)malloc(sizeof(float) * nz);
    float* sigma = (float*)malloc(sizeof(float) * ms);
    float* tau = (float*)malloc(sizeof(float) * (size_t)n);
    float* x = (float*)malloc(sizeof(float) * (size_t)n);
    float* xb = (float*)malloc(sizeof(float) * (size_t)n);
    float* y = (float*)calloc(ms, sizeof(float));
    float* axb = (float*)malloc(sizeof(float) * ms);
    float* aty = (float*)malloc(sizeof(float) * (size_t)n);
    double* rn = (double*)calloc(ms, sizeof(double));
    double* cn = (double*)calloc((size_t)n, sizeof(double));

    for (int32_t j = 0; j < n; ++j) c[j] = (float)((double)obj_sign * lp->c[j]); /* 202-204 */
    for (int32_t i = 0; i < m; ++i) { /* 206-212 */
        l_r[i] = lp->row_lower[i] <= -1e29 ? -OR_FLT_INF : (float)lp->row_lower[i];
        u_r[i] = lp->row_upper[i] >= 1e29 ? OR_FLT_INF : (float)lp->row_upper[i];
    }
    for (int32_t j = 0; j < n; ++j) { /* 214-220 */
        l_x[j] = lp->var_lower[j] <= -1e29 ? -OR_FLT_INF : (float)lp->var_lower[j];
        u_x[j] = lp->var_upper[j] >= 1e29 ? OR_FLT_INF : (float)lp->var_upper[j];
    }
    for (int32_t k = 0; k < nnz; ++k) v[k] = (float)lp->values[k]; /* 229 */

    const float alpha = 0.95f; /* 237-250 */
    for (int32_t i = 0; i < m; ++i) sigma[i] = alpha;
    for (int32_t j = 0; j < n; ++j) tau[j] = alpha;
    for (int32_t i = 0; i < m; ++i)
        for (int32_t k = lp->row_ptr[i]; k < lp->row_ptr[i + 1]; ++k) {
            const double av = fabs((double)v[k]);
            rn[i] += av;
            cn[lp->col_idx[k]] += av;
        }
    for (int32_t i = 0; i < m; ++i) if (rn[i] > 1e-12) sigma[i] = alpha / (float)rn[i];
    for (int32_t j = 0; j < n; ++j) if (cn[j] > 1e-12) tau[j] = alpha / (float)cn[j];

    for (int32_t j = 0; j < n; ++j) { /* 288-297 */
        x[j] = fmax_f(l_x[j], fmin_f(0.0f, u_x[j]));
        xb[j] = x[j];
    }

    double best = INFINITY; /* 310-314 */
    int32_t stagnation = 0, restarts = 0;
    const int32_t kRestP = 8, kMaxR = 10;
    double pv = 0.0, dv = 0.0;
    for (int32_t iter = 0; iter < max_iterations; ++iter) { /* 319-383 */
        if (m > 0) {
            or_spmv_f32(m, lp->row_ptr, lp->col_idx, v, xb, axb);
            or_dual_update(m, axb, sigma, l_r, u_r, y);
        }
        if (m > 0) or_spmv_t_f32(m, n, lp->row_ptr, lp->col_idx, v, y, aty);
        else for (int32_t j = 0; j < n; ++j) aty[j] = 0.0f;
        or_primal_update(n, aty, tau, c, l_x, u_x, x, xb);
        if ((iter + 1) % check_frequency == 0) {
            or_fp64_residuals(lp, x, y, c, v, tol, &pv, &dv);
            res->primal_residual = pv;
            res->dual_residual = dv;
            if (pv <= tol && dv <= tol) {
                res->status = 0;
                res->iterations = iter + 1;
                break;
            }
            const double combined = pv + dv;
            if (combined > tol * 10.0) {
                if (combined < best * 0.99) { best = combined; stagnation = 0; }
                else ++stagnation;
                if (stagnation >= kRestP && restarts < kMaxR) {
                    for (int32_t i = 0; i < m; ++i) y[i] = 0.0f;
                    for (int32_t j = 0; j < n; ++j) xb[j] = x[j];
                    stagnation = 0;
                    best = INFINITY;
                    ++restarts;
                }
            }
        }
    }
    if (res->status != 0) res->iterations = max_iterations; /* 392-393 */
    or_fp64_residuals(lp, x, y, c, v, tol, &pv, &dv);        /* 396-399 */
    res->primal_residual = pv;
    res->dual_residual = dv;
    res->restarts = restarts;
    double obj = lp->offset; /* 402-408 */
    for (int32_t j = 0; j < n; ++j) obj += lp->c[j] * (double)x[j];
    res->objective = obj;
    if (x_out) memcpy(x_out, x, sizeof(float) * (size_t)n);
    if (y_out && m > 0) memcpy(y_out, y, sizeof(float) * (size_t)m);

    free(c); free(l_r); free(u_r); free(l_x); free(u_x); free(v);
    free(sigma); free(tau); free(x); free(xb); free(y); free(axb); free(aty);
    free(rn); free(cn);
    return 0;
}

/* CSRMatrix::transpose, sparse_matrix.cpp:13-57: counting sort, rows
 * scattered in increasing order so each transposed row stays sorted. */
void or_transpose(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci,
                  const double* v, int32_t* t_rp, int32_t* t_ci, double* t_v) {
    for (int32_t j = 0; j <= n; ++j) t_rp[j] = 0;
    for (int32_t k = 0; k < rp[m]; ++k) ++t_rp[ci[k] + 1];
    for (int32_t j = 0; j < n; ++j) t_rp[j + 1] += t_rp[j];
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    memcpy(cur, t_rp, sizeof(int32_t) * (size_t)(n + 1));
    for (int32_t i = 0; i < m; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const int32_t d = cur[ci[k]]++;
            t_ci[d] = i;
            if (v && t_v) t_v[d] = v[k];
        }
    free(cur);
}

/* ruiz_scale, scaling.cpp:29-150: per pass, inf-norms of rows and of
 * columns (columns including |c_j|), scale by 1/sqrt(norm) where the norm
 * is positive, finite bounds only (|bound| < 1e30). */
void or_ruiz_scale(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci,
                   double* v, double* c, double* row_lower, double* row_upper,
                   double* var_lower, double* var_upper, int32_t passes,
                   double* row_scale, double* col_scale) {
    for (int32_t i = 0; i < m; ++i) row_scale[i] = 1.0;
    for (int32_t j = 0; j < n; ++j) col_scale[j] = 1.0;
    if (m == 0 || n == 0) return;
    double* rnorm = (double*)malloc(sizeof(double) * (size_t)m);
    double* cnorm = (double*)malloc(sizeof(double) * (size_t)n);
    double* rs = (double*)malloc(sizeof(double) * (size_t)m);
    double* cs = (double*)malloc(sizeof(double) * (size_t)n);
    for (int32_t p = 0; p < passes; ++p) {
        for (int32_t i = 0; i < m; ++i) {
            rnorm[i] = 0.0;
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
                if (fabs(v[k]) > rnorm[i]) rnorm[i] = fabs(v[k]);
        }
        for (int32_t j = 0; j < n; ++j) cnorm[j] = 0.0;
        for (int32_t k = 0; k < rp[m]; ++k)
            if (fabs(v[k]) > cnorm[ci[k]]) cnorm[ci[k]] = fabs(v[k]);
        for (int32_t j = 0; j < n; ++j)
            if (fabs(c[j]) > cnorm[j]) cnorm[j] = fabs(c[j]);
        for (int32_t i = 0; i < m; ++i) rs[i] = rnorm[i] > 0.0 ? 1.0 / sqrt(rnorm[i]) : 1.0;
        for (int32_t j = 0; j < n; ++j) cs[j] = cnorm[j] > 0.0 ? 1.0 / sqrt(cnorm[j]) : 1.0;
        for (int32_t i = 0; i < m; ++i)
            for (int32_t k = rp[i]; k < rp[i + 1]; ++k) v[k] *= rs[i] * cs[ci[k]];
        for (int32_t j = 0; j < n; ++j) c[j] *= cs[j];
        for (int32_t i = 0; i < m; ++i) {
            if (row_lower[i] > -OR_INF) row_lower[i] = row_lower[i] * rs[i];
            if (row_upper[i] < OR_INF) row_upper[i] = row_upper[i] * rs[i];
        }
        for (int32_t j = 0; j < n; ++j) {
            if (var_lower[j] > -OR_INF) var_lower[j] = var_lower[j] / cs[j];
            if (var_upper[j] < OR_INF) var_upper[j] = var_upper[j] / cs[j];
        }
        for (int32_t i = 0; i < m; ++i) row_scale[i] *= rs[i];
        for (int32_t j = 0; j < n; ++j) col_scale[j] *= cs[j];
    }
    free(rnorm); free(cnorm); free(rs); free(cs);
}

/* Relative KKT error of a primal-dual pair for the minimisation form of
 * the LP (c negated for a maximisation), y in the reference sign, i.e. the
 * textbook multiplier is -y and the reduced cost is c + A^T y.
 *   primal: r_i = (Ax)_i - proj_[l_i,u_i]((Ax)_i), plus x_j - proj(x_j);
 *           rel_primal = ||r||_2 / (1 + ||bhat||_2), bhat_i = largest finite
 *           |row bound| of row i (0 if none).
 *   dual:   rc = c + A^T y; lambda_j = rc_j projected on the sign the
 *           finite bounds of x_j allow (free: 0, lower only: >= 0, upper
 *           only: <= 0, boxed: any); the multiplier t = -y projected on the
 *           sign its row allows (>= 0 with only a lower row bound, <= 0
 *           with only an upper one, 0 for a free row);
 *           rel_dual = ||(rc - lambda, t - proj(t))||_2 / (1 + ||c||_2).
 *   gap:    p = c^T x; d = sum_i (t_i > 0 ? t_i l_i : t_i u_i) over the
 *           projected t + sum_j (lambda_j > 0 ? lambda_j l_j : lambda_j u_j);
 *           rel_gap = |p - d| / (1 + |p| + |d|). */
void or_relative_kkt(const or_lp* lp, const double* x, const double* y,
                     double* out) {
    const int32_t m = lp->m, n = lp->n;
    const double s = lp->sense == 1 ? -1.0 : 1.0;
    double* aty = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double rp2 = 0.0, rpinf = 0.0, b2 = 0.0, rd2 = 0.0, c2 = 0.0, p = 0.0, d = 0.0;
    for (int32_t i = 0; i < m; ++i) {
        double act = 0.0;
        for (int32_t k = lp->row_ptr[i]; k < lp->row_ptr[i + 1]; ++k) {
            act += lp->values[k] * x[lp->col_idx[k]];
            aty[lp->col_idx[k]] += lp->values[k] * y[i];
        }
        const double lo = lp->row_lower[i], hi = lp->row_upper[i];
        const int lf = lo > -OR_INF, uf = hi < OR_INF;
        double proj = act;
        if (lf && proj < lo) proj = lo;
        if (uf && proj > hi) proj = hi;
        const double r = act - proj;
        rp2 += r * r;
        if (fabs(r) > rpinf) rpinf = fabs(r);
        double bh = 0.0;
        if (lf && fabs(lo) > bh) bh = fabs(lo);
        if (uf && fabs(hi) > bh) bh = fabs(hi);
        b2 += bh * bh;
        const double t = -y[i];
        double tp = t;
        if (!lf && tp > 0.0) tp = 0.0;
        if (!uf && tp < 0.0) tp = 0.0;
        rd2 += (t - tp) * (t - tp);
        if (tp > 0.0) d += tp * lo;
        else if (tp < 0.0) d += tp * hi;
    }
    for (int32_t j = 0; j < n; ++j) {
        const double lo = lp->var_lower[j], hi = lp->var_upper[j];
        const int lf = lo > -OR_INF, uf = hi < OR_INF;
        double xp = x[j];
        if (lf && xp < lo) xp = lo;
        if (uf && xp > hi) xp = hi;
        rp2 += (x[j] - xp) * (x[j] - xp);
        const double cj = s * lp->c[j];
        c2 += cj * cj;
        p += cj * x[j];
        const double rc = cj + aty[j];
        double lam = rc;
        if (!lf && lam > 0.0) lam = 0.0;
        if (!uf && lam < 0.0) lam = 0.0;
        rd2 += (rc - lam) * (rc - lam);
        if (lam > 0.0) d += lam * lo;
        else if (lam < 0.0) d += lam * hi;
    }
    const double rel_p = sqrt(rp2) / (1.0 + sqrt(b2));
    const double rel_d = sqrt(rd2) / (1.0 + sqrt(c2));
    const double rel_g = fabs(p - d) / (1.0 + fabs(p) + fabs(d));
    double k = rel_p;
    if (rel_d > k) k = rel_d;
    if (rel_g > k) k = rel_g;
    out[0] = k; out[1] = rel_p; out[2] = rel_d; out[3] = rel_g;
    out[4] = p; out[5] = d; out[6] = rpinf;
    free(aty);
}
