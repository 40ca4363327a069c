// precision_ladder.cpp -- the precision ladder over the GPU engine; see
// include/pramaan/gpu/precision_ladder.hpp.  Rung 1 and the gate follow the
// reference controller (src/gpu/precision_ladder.cpp:57-179): the GPU x is
// returned only when the status is optimal and both reference-style
// residuals are within the threshold; otherwise the first failing condition
// names the escalation reason.
//
// PRAMAAN_ENABLE_HIP (set by the build of the GPU library) compiles the GPU
// rungs in.  Without it the controller needs no GPU code at all: it
// escalates at once with kGpuDisabled to the CPU polish, as the reference
// does when built without CUDA (src/gpu/precision_ladder.cpp:156-173).
#include "pramaan/gpu/precision_ladder.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>

#include "pramaan/gpu/pdhg_solver.hpp"

namespace pramaan {

std::string PrecisionLadderResult::escalation_reason_str() const {
    switch (escalation_reason) {
        case EscalationReason::kNone: return "none";
        case EscalationReason::kHighPrimalResidual: return "high_primal_residual";
        case EscalationReason::kHighDualResidual: return "high_dual_residual";
        case EscalationReason::kGpuIterationLimit: return "gpu_iteration_limit";
        case EscalationReason::kGpuNumericalFailure: return "gpu_numerical_failure";
        case EscalationReason::kGpuNotOptimal: return "gpu_not_optimal";
        case EscalationReason::kGpuDisabled: return "gpu_disabled";
    }
    return "unknown";
}

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

#ifdef PRAMAAN_ENABLE_HIP
// The gate: kNone when r passes, else the first reason it fails.
PrecisionLadderResult::EscalationReason gate(const SolveResult& r, double T) {
    using R = PrecisionLadderResult::EscalationReason;
    switch (r.status) {
        case SolveStatus::kOptimal:
            if (r.primal_residual > T) return R::kHighPrimalResidual;
            if (r.dual_residual > T) return R::kHighDualResidual;
            return R::kNone;
        case SolveStatus::kIterationLimit: return R::kGpuIterationLimit;
        case SolveStatus::kNumericalFailure: return R::kGpuNumericalFailure;
        default: return R::kGpuNotOptimal;
    }
}

#ifdef PRAMAAN_HIP_EXTENDED_OPTIONS
// The reference's residuals (src/gpu/pdhg_solver.cu:107-161) in FP64
// throughout: pv = largest row or bound violation of x; dv = largest
// projected reduced-cost violation / max(1, |c|inf), a column counting as
// at a bound within tol.  y in the reference sign (reduced cost c + A^T y,
// c of the minimisation form).
void fp64_residuals(const ModelIR& model, const std::vector<double>& x, const std::vector<double>& y, double tol,
                    double& pv, double& dv) {
    const int m = model.numRows(), n = model.numVars();
    const double sg = model.obj_sense == ObjSense::kMaximize ? -1.0 : 1.0;
    const std::vector<double> act = model.A.multiply(x);
    pv = 0.0;
    for (int i = 0; i < m; ++i) {
        if (act[i] < model.row_lower[i]) pv = std::max(pv, model.row_lower[i] - act[i]);
        if (act[i] > model.row_upper[i]) pv = std::max(pv, act[i] - model.row_upper[i]);
    }
    for (int j = 0; j < n; ++j) {
        if (x[j] < model.var_lower[j]) pv = std::max(pv, model.var_lower[j] - x[j]);
        if (x[j] > model.var_upper[j]) pv = std::max(pv, x[j] - model.var_upper[j]);
    }
    std::vector<double> rc(n);
    double c_norm = 1.0;
    for (int j = 0; j < n; ++j) {
        rc[j] = sg * model.obj_coeffs[j];
        c_norm = std::max(c_norm, std::abs(model.obj_coeffs[j]));
    }
    const auto& rp = model.A.rowPtr();
    const auto& ci = model.A.colIdx();
    const auto& v = model.A.values();
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) rc[ci[k]] += v[k] * y[i];
    dv = 0.0;
    for (int j = 0; j < n; ++j) {
        const bool at_lb = x[j] <= model.var_lower[j] + tol, at_ub = x[j] >= model.var_upper[j] - tol;
        double viol = 0.0;
        if (!at_lb && !at_ub) viol = std::abs(rc[j]);
        else if (at_lb && !at_ub) viol = std::max(0.0, -rc[j]);
        else if (at_ub && !at_lb) viol = std::max(0.0, rc[j]);
        dv = std::max(dv, viol);
    }
    dv /= c_norm;
}
#endif
#endif  // PRAMAAN_ENABLE_HIP

}  // namespace

PrecisionLadderResult PrecisionLadder::solve(const ModelIR& model) const {
    using Stage = PrecisionLadderResult::Stage;
    using Reason = PrecisionLadderResult::EscalationReason;
    PrecisionLadderResult result;
    const auto t_start = Clock::now();
    SolveResult best;
#ifdef PRAMAAN_ENABLE_HIP
    const double T = options_.residual_threshold;

    // Rung 1: GPU FP32 PDHG (reference algorithm).
    gpu::PdhgOptions o1;
    o1.tolerance = options_.gpu_tolerance;
    o1.max_iterations = options_.gpu_max_iterations;
    o1.check_frequency = options_.gpu_check_frequency;
    auto t = Clock::now();
    {
        gpu::PdhgSolver s1(o1);
        best = s1.solve(model);
    }
    result.gpu_elapsed_ms = ms_since(t);
    result.gpu_status = best.status;
    result.gpu_iterations = best.iterations;
    result.gpu_objective = best.objective_value;
    result.gpu_primal_residual = best.primal_residual;
    result.gpu_dual_residual = best.dual_residual;
    result.escalation_reason = gate(best, T);
    if (result.escalation_reason == Reason::kNone) {
        result.stage = Stage::kGpuFp32;
        result.escalated = false;
        result.final_result = std::move(best);
        result.total_elapsed_ms = ms_since(t_start);
        return result;
    }
    result.escalated = true;

#ifdef PRAMAAN_HIP_EXTENDED_OPTIONS
    // Rung 2 (extension): FP64-refined GPU PDHG to a relative KKT error of
    // 1e-2 x the threshold.  The gate's reference residuals are absolute
    // (pv) or count a column inside its bounds beyond the threshold as
    // interior (dv), so a relative target does not imply them: when the
    // gate fails, the rung solves again from its own (x, y) with a target
    // 100x tighter, at most kRefinedAttempts solves in all.
    if (options_.gpu_refined_rung) {
        constexpr int kRefinedAttempts = 3;
        gpu::PdhgOptions o2;
        o2.pdlp = true;
        o2.relative_kkt = true;
        o2.tolerance = std::max(1e-13, 1e-2 * std::min(T, options_.gpu_tolerance));
        o2.max_iterations = options_.refined_max_iterations;
        o2.check_frequency = 64;
        if (static_cast<int>(best.x.size()) == model.numVars()) o2.x0 = best.x;
        t = Clock::now();
        result.refined_rung_run = true;
        result.refined_iterations = 0;
        SolveResult r2;
        for (int attempt = 0; attempt < kRefinedAttempts; ++attempt) {
            gpu::PdhgSolver s2(o2);
            r2 = s2.solve(model);
            // Gate residuals of this rung in FP64 with the threshold as the
            // activity tolerance (the solver reports them at its own).
            if (static_cast<int>(r2.x.size()) == model.numVars() &&
                static_cast<int>(s2.lastDual().size()) == model.numRows())
                fp64_residuals(model, r2.x, s2.lastDual(), T, r2.primal_residual, r2.dual_residual);
            result.refined_status = r2.status;
            result.refined_primal_residual = r2.primal_residual;
            result.refined_dual_residual = r2.dual_residual;
            result.refined_rel_kkt = s2.lastRelativeKkt();
            result.refined_iterations += r2.iterations;
            if (gate(r2, T) == Reason::kNone || r2.status != SolveStatus::kOptimal || o2.tolerance <= 1e-13) break;
            o2.tolerance = std::max(1e-13, 1e-2 * o2.tolerance);
            o2.x0 = r2.x;
            o2.y0 = s2.lastDual();
        }
        result.refined_elapsed_ms = ms_since(t);
        if (gate(r2, T) == Reason::kNone) {
            result.stage = Stage::kGpuFp64Refined;
            result.final_result = std::move(r2);
            result.total_elapsed_ms = ms_since(t_start);
            return result;
        }
        if (r2.status == SolveStatus::kOptimal || best.status != SolveStatus::kOptimal) best = std::move(r2);
    }
#endif
#else
    // Built without the GPU engine: no GPU rung runs.
    result.escalated = true;
    result.escalation_reason = Reason::kGpuDisabled;
    best.status = SolveStatus::kIterationLimit;
#endif

    // Rung 3: CPU FP64 polish seeded with the GPU x.
    RevisedSimplex::Options cpu_opts = options_.cpu_options;
    if (static_cast<int>(best.x.size()) == model.numVars()) cpu_opts.primal_start_hint = best.x;
    result.stage = Stage::kCpuFp64Polish;
    const auto t_cpu = Clock::now();
    if (options_.cpu_polish) {
        result.final_result = options_.cpu_polish(model, cpu_opts);
    } else {
#ifdef PRAMAAN_LADDER_CPU_SIMPLEX
        RevisedSimplex cpu(cpu_opts);
        result.final_result = cpu.solve(model);
#else
        result.polish_unavailable = true;
        if (result.escalation_reason != Reason::kGpuDisabled)
            result.stage = result.refined_rung_run ? Stage::kGpuFp64Refined : Stage::kGpuFp32;
        result.final_result = std::move(best);
#endif
    }
    result.cpu_elapsed_ms = ms_since(t_cpu);
    result.total_elapsed_ms = ms_since(t_start);
    return result;
}

}  // namespace pramaan

// ---- C entry point (include/pramaan_ladder.h) ------------------------------
#include <cstdio>

#include "pramaan_ladder.h"

extern "C" int pm_ladder_solve(const pm_lp* lp, const pm_ladder_options* opts, pm_ladder_result* out, char* err,
                               int errlen) {
    try {
        if (!lp || !out) throw std::invalid_argument("NULL argument");
        using namespace pramaan;
        std::vector<CSRMatrix::Index> rp(lp->row_ptr, lp->row_ptr + lp->m + 1), ci(lp->col_idx, lp->col_idx + lp->nnz);
        ModelIR model(lp->sense == 1 ? ObjSense::kMaximize : ObjSense::kMinimize, lp->offset,
                      std::vector<double>(lp->c, lp->c + lp->n),
                      CSRMatrix(rp, ci, std::vector<double>(lp->values, lp->values + lp->nnz), lp->n),
                      std::vector<double>(lp->row_lower, lp->row_lower + lp->m),
                      std::vector<double>(lp->row_upper, lp->row_upper + lp->m), std::vector<std::string>(lp->m),
                      std::vector<double>(lp->var_lower, lp->var_lower + lp->n),
                      std::vector<double>(lp->var_upper, lp->var_upper + lp->n),
                      std::vector<VarType>(lp->n, VarType::kContinuous), std::vector<std::string>(lp->n));
        PrecisionLadder::Options o;
        if (opts) {
            o.residual_threshold = opts->residual_threshold;
            o.gpu_max_iterations = opts->gpu_max_iterations;
            o.gpu_check_frequency = opts->gpu_check_frequency;
            o.gpu_tolerance = opts->gpu_tolerance;
            o.gpu_refined_rung = opts->gpu_refined_rung != 0;
            o.refined_max_iterations = opts->refined_max_iterations;
        }
        const PrecisionLadderResult r = PrecisionLadder(o).solve(model);
        double* x = out->x;
        *out = pm_ladder_result{};
        out->x = x;
        out->stage = static_cast<int32_t>(r.stage);
        out->escalated = r.escalated ? 1 : 0;
        out->escalation_reason = static_cast<int32_t>(r.escalation_reason);
        out->gpu_status = static_cast<int32_t>(r.gpu_status);
        out->gpu_primal_residual = r.gpu_primal_residual;
        out->gpu_dual_residual = r.gpu_dual_residual;
        out->gpu_objective = r.gpu_objective;
        out->gpu_iterations = r.gpu_iterations;
        out->refined_rung_run = r.refined_rung_run ? 1 : 0;
        out->refined_status = static_cast<int32_t>(r.refined_status);
        out->refined_iterations = r.refined_iterations;
        out->refined_rel_kkt = r.refined_rel_kkt;
        out->polish_unavailable = r.polish_unavailable ? 1 : 0;
        out->final_status = static_cast<int32_t>(r.final_result.status);
        out->final_objective = r.final_result.objective_value;
        out->final_primal_residual = r.final_result.primal_residual;
        out->final_dual_residual = r.final_result.dual_residual;
        if (x && static_cast<int>(r.final_result.x.size()) == lp->n)
            std::copy(r.final_result.x.begin(), r.final_result.x.end(), x);
        return 0;
    } catch (const std::exception& e) {
        if (err && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", e.what());
        return 1;
    }
}
