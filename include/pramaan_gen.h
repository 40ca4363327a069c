/*
 * pramaan_gen.h -- synthetic LP workloads with a known optimum.
 *
 * No reference counterpart: the reference builds its benchmark LPs inside
 * its harnesses (benchmark/rolling_horizon_bench.cpp:131-233).  These
 * generators produce the workloads the engine is measured on, each with a
 * primal-dual pair (x*, y*) that satisfies the KKT conditions exactly, so
 * the optimum p* = c^T x* is known in closed form.
 *
 * Construction (minimisation, y in the textbook sign):
 *   A: for each column j, k_j distinct rows drawn uniformly without
 *      replacement (k_j uniform in [k_min, k_max], capped at m), values
 *      N(0,1) * (1 + U(-noise, noise)) rounded to FP32.
 *   x*: a random floor(support * n) columns get U(0,1), the rest 0;
 *      bounds 0 <= x <= var_upper.
 *   rows: a random round(eq_frac * m) rows are equalities b_i = (A x*)_i
 *      with y*_i ~ N(0,1); the others are "<=" rows with
 *      b_i = (A x*)_i + U(0,1) and y*_i = 0 (inactive).
 *   c = A^T y* + z with z_j = 0 where x*_j > 0 and U(0,1) elsewhere.
 * The engine's dual (reference sign) converges to -y*.
 */
#ifndef PRAMAAN_GEN_H
#define PRAMAAN_GEN_H

#include <stdint.h>

#include "pramaan_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_gen_params {
    int32_t m, n;
    int32_t k_min, k_max;   /* nonzeros per column                        */
    double eq_frac;         /* fraction of equality rows (1.0: all)        */
    double noise;           /* relative value noise (0.05: +-5 %)          */
    double support;         /* fraction of columns with x*_j > 0 (0.3)     */
    double var_upper;       /* upper bound of every x_j (10)               */
    uint64_t seed;
} pm_gen_params;

typedef struct pm_gen_lp pm_gen_lp;

/* Parameters of the named workloads:
 *  0: m=2,000 n=5,000, 8 nnz/col, all equality rows (config 0)
 *  1: m=500,000 n=1,000,000, 12 nnz/col, 60 % equality, 5 % noise
 *  2: 1/10 of workload 1: m=50,000 n=100,000
 *  3: one instance of the batch workload: m log-uniform in [50, 5,000]
 *     drawn from the seed, n = 2.5 m, 4-16 nnz/col, all equality rows */
int pm_gen_params_named(int32_t which, uint64_t seed, pm_gen_params* out);
int pm_gen_synthetic(const pm_gen_params* p, pm_gen_lp** out);
/* Views into the generated arrays (valid until pm_gen_free). */
int pm_gen_view(const pm_gen_lp* g, pm_lp* lp, const double** x_star,
                const double** y_star, double* p_star);
void pm_gen_free(pm_gen_lp* g);

/* ---- rolling-horizon blending LPs --------------------------------------
 * A T-period blending plan (24 components, 15 products, 8 quality
 * attributes, 10 processing units).  A period has 400 columns and 215
 * rows: a 200-row block (120 maximum- and 21 minimum-quality rows, 24
 * supply, 10 capacity, 10 market rows, then 15 demand rows after the
 * balance rows) and 15 inventory-balance rows that link it to the period
 * before.  Window w is the LP of periods [w, w + window) with the
 * exogenous opening stock of period w; column tau*400 + j and row
 * tau*215 + i of a window are column j / row i of its period tau.  The
 * default horizon is 523 periods, i.e. 500 windows of 24.  demand_noise
 * and blend_noise (with perturb_seed) scale every demand and blend
 * coefficient by U(1 - noise, 1 + noise).  Windows carry no closed-form
 * optimum (x_star / y_star views are empty, p_star is NaN). */
typedef struct pm_rh_params {
    int32_t periods_total;
    int32_t window;
    double demand_noise;
    double blend_noise;
    uint64_t seed;
    uint64_t perturb_seed;
} pm_rh_params;
typedef struct pm_rh pm_rh;
int pm_rh_params_default(pm_rh_params* p);
int pm_rh_create(const pm_rh_params* p, pm_rh** out);
int pm_rh_window_size(const pm_rh* h, int32_t* rows_per_period,
                      int32_t* cols_per_period, int32_t* windows);
int pm_rh_window(const pm_rh* h, int32_t start, pm_gen_lp** out);
void pm_rh_free(pm_rh* h);

#ifdef __cplusplus
}
#endif
#endif
