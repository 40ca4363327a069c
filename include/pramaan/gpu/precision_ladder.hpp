// precision_ladder.hpp -- GPU fast pass, acceptance gate, escalation.
//
// Drop-in for the reference's PrecisionLadder
// (include/pramaan/gpu/precision_ladder.hpp:93-199,
// src/gpu/precision_ladder.cpp:57-179): the same result and option types,
// the same gate (GPU status optimal and the reference-style primal and dual
// residuals both <= residual_threshold) and the same escalation reasons.
//
// Rungs:
//   1. GPU FP32 PDHG, the reference algorithm with the gpu_* options.
//   2. (extension, off by default) GPU PDHG with FP64 iterative refinement
//      (reflected Halpern, relative-KKT termination at 1e-2 x the
//      threshold), warm-started from rung 1's x, judged by the same gate on
//      the reference residuals recomputed in FP64 with the threshold as the
//      activity tolerance; a failed gate re-solves from the rung's own
//      (x, y) at a 100x tighter target, at most three solves in all.  It
//      turns a GPU result into the answer where the reference would fall
//      back to the CPU.
//   Built without PRAMAAN_ENABLE_HIP, no GPU rung exists: the ladder
//   escalates at once with kGpuDisabled, as the reference built without
//   CUDA does.
//   3. CPU FP64 polish with the best GPU x as primal_start_hint:
//      Options::cpu_polish when set; otherwise RevisedSimplex when the
//      library is built inside the PRAMAAN tree (PRAMAAN_LADDER_CPU_SIMPLEX);
//      otherwise no CPU solver is linked and the best GPU result is
//      returned with polish_unavailable = true.
#pragma once

#include <functional>
#include <limits>
#include <string>
#include <vector>

#include "pramaan/ir.hpp"
#include "pramaan/simplex.hpp"

namespace pramaan {

struct PrecisionLadderResult {
    enum class Stage {
        kGpuFp32,         // rung 1 met the threshold
        kCpuFp64Polish,   // the CPU FP64 solver produced the final result
        kGpuFp64Refined,  // rung 2 (extension) met the threshold
    };
    enum class EscalationReason {
        kNone,
        kHighPrimalResidual,
        kHighDualResidual,
        kGpuIterationLimit,
        kGpuNumericalFailure,
        kGpuNotOptimal,
        kGpuDisabled,
    };

    SolveResult final_result;
    Stage stage = Stage::kCpuFp64Polish;

    // rung 1
    SolveStatus gpu_status = SolveStatus::kIterationLimit;
    double gpu_primal_residual = std::numeric_limits<double>::infinity();
    double gpu_dual_residual = std::numeric_limits<double>::infinity();
    double gpu_objective = 0.0;
    int gpu_iterations = 0;
    double gpu_elapsed_ms = 0.0;
    double cpu_elapsed_ms = 0.0;

    bool escalated = false;
    EscalationReason escalation_reason = EscalationReason::kNone;
    double total_elapsed_ms = 0.0;

    // ---- extensions ----
    bool refined_rung_run = false;
    SolveStatus refined_status = SolveStatus::kIterationLimit;
    double refined_primal_residual = std::numeric_limits<double>::infinity();
    double refined_dual_residual = std::numeric_limits<double>::infinity();
    double refined_rel_kkt = std::numeric_limits<double>::infinity();
    int refined_iterations = 0;
    double refined_elapsed_ms = 0.0;
    bool polish_unavailable = false;  // rejected, and no CPU solver was linked

    std::string escalation_reason_str() const;
};

class PrecisionLadder {
public:
    struct Options {
        double residual_threshold = 1e-6;
        int gpu_max_iterations = 100000;
        int gpu_check_frequency = 100;
        double gpu_tolerance = 1e-6;
        RevisedSimplex::Options cpu_options;
        // ---- extensions ----
        bool gpu_refined_rung = false;
        int refined_max_iterations = 200000;
        std::function<SolveResult(const ModelIR&, const RevisedSimplex::Options&)> cpu_polish;
    };

    PrecisionLadder() = default;
    explicit PrecisionLadder(Options options) : options_(std::move(options)) {}

    PrecisionLadderResult solve(const ModelIR& model) const;
    const Options& options() const noexcept { return options_; }

private:
    Options options_{};
};

}  // namespace pramaan
