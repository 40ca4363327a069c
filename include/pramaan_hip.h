/*
 * pramaan_hip.h -- C-ABI of the MI355X (gfx950) PDHG LP engine.
 *
 * This is the drop-in boundary for PRAMAAN's GPU PDHG path.  Every entry
 * point takes plain pointers and sizes (no C++ or torch types), returns an
 * int status (PM_OK on success) and never throws; the message of the last
 * failure on the calling thread is available from pm_last_error().
 *
 * The C++ solver interface (pramaan::gpu::PdhgSolver, pramaan::Backend) in
 * include/pramaan/ is a thin header-level wrapper over these functions.
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   pm_pdhg_solve            <- gpu::PdhgSolver::solve
 *                               include/pramaan/gpu/pdhg_solver.hpp:15-24,
 *                               src/gpu/pdhg_solver.cu:169-411
 *   pm_pdhg_options_default  <- gpu::PdhgOptions defaults
 *                               include/pramaan/gpu/pdhg_solver.hpp:9-13
 *   pm_launch_spmv           <- gpu::launch_spmv   src/gpu/pdhg.cu:156-196
 *   pm_launch_spmv_t         <- gpu::launch_spmv_t src/gpu/pdhg.cu:198-241
 *   pm_launch_pdhg_iteration <- gpu::launch_pdhg_iteration src/gpu/pdhg.cu:77-146
 *   pm_ruiz_scale            <- pramaan::ruiz_scale
 *                               include/pramaan/presolve.hpp:45,
 *                               src/presolve/scaling.cpp:29-150
 *   pm_last_error            <- the message CUDA_CHECK puts in the
 *                               std::runtime_error it throws
 *                               src/gpu/cuda_utils.cuh:8-15
 * The device-resident context (pm_ctx_*) and the batched solver
 * (pm_batch_solve) have no reference counterpart: the reference solves one
 * LP per call and re-uploads it every time (pdhg_solver.cu:252-297).
 */
#ifndef PRAMAAN_HIP_H
#define PRAMAAN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Infinite bound sentinel, ModelIR's kInfinity (include/pramaan/ir.hpp:47). */
#define PM_INFINITY_VALUE 1.0e30

/* ---- return codes ------------------------------------------------------ */
#define PM_OK                 0
#define PM_ERR_INVALID_ARG    1 /* bad sizes, null pointers, inconsistent CSR */
#define PM_ERR_HIP            2 /* a HIP runtime call failed               */
#define PM_ERR_NO_DEVICE      3 /* no usable gfx950 device                  */
#define PM_ERR_OUT_OF_MEMORY  4
#define PM_ERR_INTERNAL       5

/* ---- SolveStatus (include/pramaan/simplex.hpp:26-34), same order ------- */
#define PM_SOLVE_OPTIMAL              0
#define PM_SOLVE_INFEASIBLE           1
#define PM_SOLVE_UNBOUNDED            2
#define PM_SOLVE_ITERATION_LIMIT      3
#define PM_SOLVE_NUMERICAL_FAILURE    4
#define PM_SOLVE_WARM_START_REJECTED  5

/* ---- algorithm --------------------------------------------------------- */
/* PM_ALG_REFERENCE: the reference PDHG (pdhg_solver.cu:169-411): diagonal
 *   Pock-Chambolle steps 0.95/sum|A| (237-250), FP32 iterations, FP64
 *   residual check every check_frequency iterations, stagnation restarts
 *   (307-381).  Same statuses, iteration count and residual definitions.
 * PM_ALG_PDLP: restarted PDHG on the Ruiz + Pock-Chambolle scaled LP with a
 *   power-iteration step size, primal-weight updates, KKT-driven adaptive
 *   restarts and an FP64 iterative-refinement outer loop.  Terminates on
 *   the FP64 relative KKT error of the original LP. */
#define PM_ALG_REFERENCE 0
#define PM_ALG_PDLP      1

/* ---- termination test ---------------------------------------------------
 * PM_ALG_REFERENCE accepts PM_TERM_REFERENCE only.  PM_ALG_PDLP under
 * PM_TERM_REFERENCE stops when the relative KKT error and the reference's
 * pv and dv all meet the tolerance (a check whose relative error meets it
 * while pv or dv do not tightens the relative target tenfold).  Other
 * pairs fail with PM_ERR_INVALID_ARG. */
#define PM_TERM_REFERENCE    0 /* pv <= tol && dv <= tol (pdhg_solver.cu:356) */
#define PM_TERM_RELATIVE_KKT 1 /* max(rel primal, rel dual, rel gap) <= tol   */
/* Restart metric of PM_ALG_PDLP (pm_pdhg_options.restart_metric). */
#define PM_RESTART_KKT 0 /* the FP64 relative KKT error of the candidate   */
#define PM_RESTART_GAP 1 /* PDLP's normalized duality gap of the candidate */
/* Step rule of PM_ALG_PDLP (pm_pdhg_options.step_rule). */
#define PM_STEP_HALPERN 0  /* reflected Halpern, constant step 0.99/||A||_2  */
#define PM_STEP_ADAPTIVE 1 /* plain PDHG with PDLP's adaptive step size      */

/* ---- SpMV variants ------------------------------------------------------ */
#define PM_SPMV_AUTO        0  /* autotuned per matrix                        */
#define PM_SPMV_SEQUENTIAL  1  /* one lane per row, sequential order, no FMA:
                                  bitwise equal to a sequential CPU loop     */
#define PM_SPMV_SCALAR      2  /* one lane per row, FMA                       */
#define PM_SPMV_VECTOR2     3  /* 2 lanes per row, wavefront-shuffle reduce  */
#define PM_SPMV_VECTOR4     4
#define PM_SPMV_VECTOR8     5
#define PM_SPMV_VECTOR16    6
#define PM_SPMV_VECTOR32    7
#define PM_SPMV_VECTOR64    8
#define PM_SPMV_STREAM      9  /* LDS-staged row segments (CSR-stream), batch
                                  engine                                     */
#define PM_SPMV_SLICED      10 /* sliced ELL: 64-row slices, rows sorted by
                                  length within a window, one lane per row,
                                  every matrix load coalesced.  Batch engine:
                                  16-bit column indexes inside each group's
                                  column window.  Single-LP PDLP engine: both
                                  products of the iteration over 32-bit
                                  copies of A and A^T (windows of 1024 rows);
                                  AUTO times them beside the row kernels on
                                  LPs of at least 2^18 nonzeros.             */
#define PM_SPMV_STAGED      11 /* sliced ELL with the gathered vector of each
                                  LP staged in LDS (batch engine)            */
#define PM_SPMV_COLBLOCK    12 /* single-LP PDLP engine: A x-bar in column
                                  blocks (x-bar cut in 2, in slabs of at most
                                  1 MB or of at most 512 KB, whichever times
                                  fastest), one launch per block gathering
                                  from an L2-resident slab, row sums
                                  accumulated across blocks; A^T y and the
                                  checks autotuned as AUTO.  AUTO times it
                                  beside the row kernels on LPs with
                                  n >= 2^19 and nnz >= 2^22.              */
#define PM_SPMV_MERGE       13 /* merge-path CSR (the product split evenly
                                  over rows + nonzeros, whatever the row
                                  lengths; a row of more than 128 entries
                                  split over tiles is finished by a second
                                  pass that adds its parts in order and
                                  applies the fused update).  Single-LP PDLP
                                  engine: both
                                  products of the iteration; AUTO times it
                                  beside the row kernels on LPs of at least
                                  2^18 nonzeros.  Also pm_launch_spmv.     */
#define PM_SPMV_NUM_VARIANTS 14

/* ---- the LP, in the layout of ModelIR (include/pramaan/ir.hpp:62-119) ---
 * optimize sense: c^T x + offset  s.t.  row_lower <= A x <= row_upper,
 *                                       var_lower <= x   <= var_upper.
 * Infinite bounds are +-1e30 (ir.hpp:47).  A is CSR with int32 indices,
 * column indices sorted and unique within a row (sparse_matrix.hpp:28-35). */
typedef struct pm_lp {
    int32_t m, n, nnz;
    int32_t sense;            /* 0 = minimize, 1 = maximize                  */
    double offset;
    const double* c;          /* [n] */
    const int32_t* row_ptr;   /* [m+1] */
    const int32_t* col_idx;   /* [nnz] */
    const double* values;     /* [nnz] */
    const double* row_lower;  /* [m] */
    const double* row_upper;  /* [m] */
    const double* var_lower;  /* [n] */
    const double* var_upper;  /* [n] */
} pm_lp;

/* ---- options: PdhgOptions (pdhg_solver.hpp:9-13) plus extensions -------
 * Defaults (pm_pdhg_options_default) equal the reference's defaults, so a
 * caller that only sets the three reference fields gets the reference
 * algorithm. */
typedef struct pm_pdhg_options {
    double  tolerance;         /* 1e-6                                       */
    int32_t max_iterations;    /* 100000                                     */
    int32_t check_frequency;   /* 100                                        */
    int32_t algorithm;         /* PM_ALG_REFERENCE                           */
    int32_t termination;       /* PM_TERM_REFERENCE                          */
    int32_t spmv_variant;      /* PM_SPMV_AUTO                               */
    int32_t ruiz_iterations;   /* PDLP: Ruiz passes (10)                     */
    int32_t max_refinements;   /* PDLP: FP64 refinement rounds (8)           */
    int32_t device;            /* HIP device ordinal (0)                     */
    int32_t verbose;           /* 0: silent; 1: one line per check to stderr */
    double  primal_weight;     /* PDLP: initial primal weight, <= 0 = auto   */
    const double* x0;          /* optional warm start [n], original space
                                  (PM_ALG_PDLP only: the reference
                                  algorithm rejects a warm start)          */
    const double* y0;          /* optional warm start [m], reference sign    */
    int32_t restart_metric;    /* PDLP: what the adaptive restart rule
                                  compares, PM_RESTART_KKT (0, default) or
                                  PM_RESTART_GAP (single-LP engine only;
                                  the batch engine rejects it)              */
    int32_t step_rule;         /* PDLP: PM_STEP_HALPERN (0, default) or
                                  PM_STEP_ADAPTIVE (single-LP engine only;
                                  the batch engine rejects it)              */
    int32_t feasibility_polishing; /* PDLP: 1 (default) = once the gap is small,
                                  solve the primal and the dual feasibility
                                  problems from the candidate and accept
                                  the pair when it meets the tolerance;
                                  0 = off                                  */
} pm_pdhg_options;

/* ---- result: SolveResult (simplex.hpp:46-76) plus PDHG diagnostics ------ */
typedef struct pm_pdhg_result {
    int32_t status;            /* PM_SOLVE_*                                 */
    int32_t iterations;        /* iter+1 when optimal, else max_iterations   */
    int32_t restarts;
    int32_t refinements;
    double  objective;         /* c^T x + offset in the model's sense        */
    double  primal_residual;   /* reference pv (pdhg_solver.cu:124-134)      */
    double  dual_residual;     /* reference dv (pdhg_solver.cu:136-157)      */
    double  rel_kkt;           /* max of the three relative errors below     */
    double  rel_primal;        /* ||Ax - proj(Ax)||_2 / (1 + ||b||_2)        */
    double  rel_dual;          /* ||c + A^T y - lambda||_2 / (1 + ||c||_2)   */
    double  rel_gap;           /* |p - d| / (1 + |p| + |d|)                  */
    double  primal_objective;  /* minimization form, without offset          */
    double  dual_objective;
    double  primal_weight;
    double  device_seconds;    /* wall time of the device solve loop         */
    double* x;                 /* caller-owned [n] or NULL.  PM_ALG_PDLP
                                  without convergence (iteration limit,
                                  numerical failure): the checked candidate
                                  of lowest rel_kkt, which every field here
                                  describes; PM_ALG_REFERENCE: the last
                                  iterate, as the reference returns it     */
    double* y;                 /* caller-owned [m] or NULL (reference sign:
                                  reduced cost is c + A^T y)                 */
    double* row_activity;      /* caller-owned [m] or NULL: A x in FP64      */
    int32_t polish_attempts;   /* PDLP feasibility polishes tried (the last
                                  accepted when it ended the solve)        */
} pm_pdhg_result;

typedef struct pm_ctx pm_ctx;

/* ---- library ------------------------------------------------------------ */
const char* pm_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* pm_last_error(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int pm_device_count(int32_t* count);

/* ---- the solver --------------------------------------------------------- */
void pm_pdhg_options_default(pm_pdhg_options* opts);
/* Solves one LP.  The LP is copied to the device; res->x/y/row_activity are
 * filled when non-NULL.  Replaces PdhgSolver::solve (pdhg_solver.cu:169). */
int pm_pdhg_solve(const pm_lp* lp, const pm_pdhg_options* opts,
                  pm_pdhg_result* res);

/* ---- test-facing kernels (host vectors in, host vectors out) ----------- */
/* y[m] = A x, A in CSR (row_ptr[m+1], col_idx/values[nnz]), x[n].
 * variant: PM_SPMV_*; PM_SPMV_AUTO picks by average row length. */
int pm_launch_spmv(const int32_t* row_ptr, const int32_t* col_idx,
                   const float* values, int32_t nnz, const float* x,
                   float* y, int32_t m, int32_t n, int32_t variant);
/* y[n] = A^T x, x[m]; computed as a row-parallel SpMV over the transpose
 * CSR (deterministic, no atomics; the reference scatters with atomicAdd). */
int pm_launch_spmv_t(const int32_t* row_ptr, const int32_t* col_idx,
                     const float* values, int32_t nnz, const float* x,
                     float* y, int32_t m, int32_t n, int32_t variant);
/* One simplified PDHG step on dense vectors (pdhg.cu:42-66):
 * x <- max(0, x - tau*gradient) (n), y <- y + sigma*(ax_bar - b) (m). */
int pm_launch_pdhg_iteration(float* x, float* y, const float* gradient,
                             const float* ax_bar, const float* b, int32_t n,
                             int32_t m, float tau, float sigma);

/* ---- Ruiz equilibration ------------------------------------------------ */
/* `passes` passes over the m x n LP (CSR row_ptr[m+1], col_idx/values[nnz]):
 * each takes the infinity norm of every row and of every column (a column's
 * including |c_j|) and scales rows and columns by 1/sqrt(norm) (1 for a
 * zero norm).  In place: values, c, the finite row bounds (times the row
 * factor) and the finite variable bounds (divided by the column factor;
 * |bound| >= PM_INFINITY_VALUE is absent).  row_scale[m], col_scale[n]
 * receive the cumulative factors: A_scaled = diag(row_scale) A
 * diag(col_scale).  With m = 0, n = 0 or passes = 0 nothing changes and the
 * factors are 1.  Host arrays in and out; the passes run on device 0, with
 * the reference's FP64 expressions, so the results equal its host loops bit
 * for bit. */
int pm_ruiz_scale(int32_t m, int32_t n, const int32_t* row_ptr,
                  const int32_t* col_idx, double* values, double* c,
                  double* row_lower, double* row_upper, double* var_lower,
                  double* var_upper, int32_t passes, double* row_scale,
                  double* col_scale);

/* pm_ruiz_scale without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * array is caller-owned DEVICE memory on the current device, with the
 * meaning and results of pm_ruiz_scale (bit for bit): row_ptr[m+1],
 * col_idx/values[nnz] (values scaled in place), c, bounds in place,
 * row_scale[m], col_scale[n] written.  The per-pass factors and column
 * norms live in a fixed device array of the library (32 MiB per device,
 * held until the process ends), so m + 2n may not exceed 4,194,304: larger
 * LPs fail with PM_ERR_INVALID_ARG before any launch.  A row pointer or
 * column index outside the sizes given fails with PM_ERR_INVALID_ARG (the
 * arrays are then partly scaled).  Runs on `stream` (a hipStream_t; NULL =
 * the null stream) and returns when the work is done.  Calls that overlap
 * on one device share the array and race: serialise them. */
int pm_ruiz_scale_noalloc(int32_t m, int32_t n, int32_t nnz,
                          const int32_t* row_ptr, const int32_t* col_idx,
                          double* values, double* c, double* row_lower,
                          double* row_upper, double* var_lower,
                          double* var_upper, int32_t passes,
                          double* row_scale, double* col_scale, void* stream);

/* ---- device-resident engine -------------------------------------------- */
/* A context owns one device stream and the device copy of one LP (both CSR
 * orientations, scaled FP32 working data and FP64 originals).  A context is
 * used by one thread at a time. */
int pm_ctx_create(int32_t device, pm_ctx** out);
int pm_ctx_destroy(pm_ctx* ctx);
/* Uploads and prepares the LP (canonicalisation, scaling, transpose,
 * step sizes, SpMV autotuning). */
int pm_ctx_load(pm_ctx* ctx, const pm_lp* lp, const pm_pdhg_options* opts);
/* Solves the loaded LP from a cold start (or opts->x0/y0 given at load). */
int pm_ctx_solve(pm_ctx* ctx, pm_pdhg_result* res);
/* PM_ALG_PDLP on an LP below 2^18 nonzeros replays each full chunk of
 * check_frequency iterations as one captured HIP graph (default 1).  use_graph = 0 issues the same launches
 * directly, with identical results; for profilers that do not follow graph
 * launches.  No counterpart in the reference. */
int pm_ctx_set_graph(pm_ctx* ctx, int32_t use_graph);

/* Kernel ids for pm_ctx_time_kernel. */
#define PM_KERNEL_SPMV_A      0 /* A x-bar with the fused dual update       */
#define PM_KERNEL_SPMV_AT     1 /* A^T y with the fused primal update       */
#define PM_KERNEL_ITERATION   2 /* both, i.e. one PDHG iteration            */
/* Launches kernel `which` `reps` times back to back on the context stream
 * between two HIP events and reports the mean duration per launch and the
 * algorithmic bytes one launch moves (each array touched once). */
int pm_ctx_time_kernel(pm_ctx* ctx, int32_t which, int32_t reps,
                       double* avg_ms, double* alg_bytes);
/* Row and column scaling factors of the loaded LP (arrays of m and n):
 * D_r, D_c of PM_ALG_PDLP (Ruiz infinity-norm passes on A, then
 * Pock-Chambolle 1-norms, computed on the device; the reference applies
 * its Ruiz scaling on the host, src/presolve/scaling.cpp:29-150), all ones
 * for PM_ALG_REFERENCE. */
int pm_ctx_scaling(pm_ctx* ctx, double* row_scale, double* col_scale);
/* SpMV variant the context chose for A and A^T (PM_SPMV_*). */
int pm_ctx_spmv_variants(pm_ctx* ctx, int32_t* variant_a,
                         int32_t* variant_at);

/* ---- batched solver ------------------------------------------------------
 * Many independent LPs, resident on the device as one block-diagonal LP,
 * each running the PM_ALG_PDLP algorithm (power iteration, refinement
 * rounds, reflected Halpern iterations, FP64 KKT checks, restarts)
 * entirely on the device: every iteration is two grid-wide launches over
 * all LPs still running; the checks' FP64 products are grid-wide too, and
 * the per-LP decisions run one workgroup per LP.  The launch sequence is
 * captured once in a HIP graph (a second one for the tail, when few LPs
 * remain) and replayed by every pm_batch_run.  Only opts->tolerance,
 * max_iterations, check_frequency, max_refinements, ruiz_iterations and
 * spmv_variant are used; the algorithm is always PM_ALG_PDLP with
 * PM_TERM_RELATIVE_KKT. */
typedef struct pm_batch pm_batch;
typedef struct pm_batch_stats {
    double device_seconds;      /* wall time of the launch                   */
    double* instance_seconds;   /* caller-owned [count] or NULL: per-LP solve
                                   time measured on the device clock         */
} pm_batch_stats;
int pm_batch_create(int32_t device, int32_t count, const pm_lp* lps,
                    const pm_pdhg_options* opts, pm_batch** out);
/* Warm start of LP i for the following runs (NULL clears it); original
 * space, y in the reference sign.  Copied at the call. */
int pm_batch_set_warm_start(pm_batch* b, int32_t i, const double* x0,
                            const double* y0);
/* Initial primal weight of LP i for the following runs (the weight of the
 * Ruiz / Pock-Chambolle scaled problem, as pm_pdhg_result.primal_weight
 * reports it); omega <= 0 restores the default ||c_hat||_2 / ||b_hat||_2. */
int pm_batch_set_primal_weight(pm_batch* b, int32_t i, double omega);
/* Row and column scaling factors of LP i (Ruiz inf-norm passes then
 * Pock-Chambolle 1-norms, computed on the device for the whole batch; the
 * reference scales on the host, src/presolve/scaling.cpp:29-150);
 * row_scale[m], col_scale[n]. */
int pm_batch_scaling(pm_batch* b, int32_t i, double* row_scale, double* col_scale);
/* Solves every LP of the batch; results[i].x/.y filled when non-NULL.
 * A warm start is judged first: its FP64 relative KKT error is compared
 * with that of the cold start (x = 0 clipped to the bounds, y = 0), and
 * the LP starts from whichever is smaller. */
int pm_batch_run(pm_batch* b, pm_pdhg_result* results, pm_batch_stats* stats);
int pm_batch_destroy(pm_batch* b);
/* After a run, per LP (arrays of count, either may be NULL): accepted = 1
 * if its warm start was used, 0 if the cold start replaced it, -1 without
 * a warm start; start_rel_kkt = FP64 relative KKT error of the starting
 * point used (-1 without a warm start). */
int pm_batch_warm_start_verdicts(pm_batch* b, int32_t* accepted,
                                 double* start_rel_kkt);
/* use_graph = 0 issues the same launches directly on the stream instead of
 * replaying the captured graph (for profilers that do not follow graph
 * launches); results are identical.  Default 1. */
int pm_batch_set_graph(pm_batch* b, int32_t use_graph);
/* Times the grid-wide iterate kernel `which` (PM_KERNEL_SPMV_A: A xe with
 * the fused dual update, PM_KERNEL_SPMV_AT: A^T y with the fused primal
 * update) with every LP of the batch running: `reps` launches back to back
 * between two HIP events on the batch stream.  Reports the mean duration
 * per launch and the algorithmic bytes of one launch (the single-LP
 * formulas of pm_ctx_time_kernel summed over the LPs).  Leaves the iterates
 * changed; the next pm_batch_run starts afresh. */
int pm_batch_time_kernel(pm_batch* b, int32_t which, int32_t reps,
                         double* avg_ms, double* alg_bytes);
/* Iterate kernels used for A and A^T while many LPs run, as PM_SPMV_*
 * (SCALAR / VECTORx: lanes per row; STREAM; SLICED; STAGED):
 * opts->spmv_variant when it names one of those, else autotuned at
 * pm_batch_create over the whole batch. */
int pm_batch_spmv_variants(pm_batch* b, int32_t* variant_a,
                           int32_t* variant_at);
/* Size classes of the iterate launches (no counterpart in the reference;
 * for callers reporting the batch's kernel choice): the LPs fall in
 * PM_BATCH_SIZE_CLASSES classes by row count (m < 256, m < 1024, larger).
 * *launches = 1 when one launch per half iteration covers every LP (then
 * pm_batch_spmv_variants gives its kernels), PM_BATCH_SIZE_CLASSES when
 * the autotuner found one launch per class faster; lps[c] = LPs of class
 * c, variant_a[c] / variant_at[c] = that class's kernels (PM_SPMV_*).
 * Any output may be NULL. */
#define PM_BATCH_SIZE_CLASSES 3
int pm_batch_size_classes(pm_batch* b, int32_t* launches, int32_t* lps,
                          int32_t* variant_a, int32_t* variant_at);
/* Overrides the autotuner's choice (no counterpart in the reference; for
 * tests and tuning): launches = 1 issues one launch per half iteration over
 * every LP, PM_BATCH_SIZE_CLASSES one per non-empty size class with that
 * class's kernels (the whole batch's kernels when nothing was tuned per
 * class).  The next pm_batch_run captures its graphs again. */
int pm_batch_set_size_classes(pm_batch* b, int32_t launches);
/* create + run + destroy; x0/y0 may be NULL or arrays of count pointers. */
int pm_batch_solve(int32_t count, const pm_lp* lps, const double* const* x0,
                   const double* const* y0, const pm_pdhg_options* opts,
                   pm_pdhg_result* results, pm_batch_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* PRAMAAN_HIP_H */
