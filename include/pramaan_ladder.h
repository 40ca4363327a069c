/* pramaan_ladder.h -- C entry point of the precision ladder
 * (include/pramaan/gpu/precision_ladder.hpp) in libpramaan_cxx.so.
 *
 * Replaces a call of the reference's PrecisionLadder::solve
 * (src/gpu/precision_ladder.cpp:57-179) for a pm_lp in memory.  Built
 * outside the PRAMAAN tree the library links no CPU simplex: a result the
 * GPU rungs cannot get through the gate comes back with
 * polish_unavailable = 1 and the best GPU point. */
#ifndef PRAMAAN_LADDER_H
#define PRAMAAN_LADDER_H

#include "pramaan_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_ladder_options {
    double residual_threshold;     /* gate: status optimal, pv and dv <= this   */
    int32_t gpu_max_iterations;    /* rung 1 (reference algorithm)             */
    int32_t gpu_check_frequency;
    double gpu_tolerance;
    int32_t gpu_refined_rung;      /* 1: run rung 2 (FP64-refined PDHG)         */
    int32_t refined_max_iterations;
} pm_ladder_options;

/* Stage: 0 rung 1 accepted, 1 CPU polish, 2 rung 2 accepted.  Escalation
 * reason: 0 none, 1 high primal residual, 2 high dual residual, 3 GPU
 * iteration limit, 4 GPU numerical failure, 5 GPU not optimal, 6 GPU
 * disabled.  Statuses are PM_SOLVE_*. */
typedef struct pm_ladder_result {
    int32_t stage, escalated, escalation_reason;
    int32_t gpu_status, gpu_iterations;
    double gpu_primal_residual, gpu_dual_residual, gpu_objective;
    int32_t refined_rung_run, refined_status, refined_iterations;
    double refined_rel_kkt;
    int32_t polish_unavailable;
    int32_t final_status;
    double final_objective, final_primal_residual, final_dual_residual;
    double* x;                     /* caller-owned [n] or NULL: final x         */
} pm_ladder_result;

/* Defaults are the reference's: threshold 1e-6, 100000 iterations, check
 * every 100, tolerance 1e-6, no rung 2 (200000 iterations when enabled).
 * Returns 0, or 1 with the reason in err. */
int pm_ladder_solve(const pm_lp* lp, const pm_ladder_options* opts,
                    pm_ladder_result* out, char* err, int errlen);

#ifdef __cplusplus
}
#endif

#endif /* PRAMAAN_LADDER_H */
