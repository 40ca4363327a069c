// pdhg_solver.cpp -- gpu::PdhgSolver over the engine's C-ABI.
//
// Compiles against this project's include/pramaan headers and, unchanged,
// against the reference's (include/pramaan/gpu/pdhg_solver.hpp:9-24), where
// PdhgOptions has only its three reference fields: everything beyond them
// is guarded by PRAMAAN_HIP_EXTENDED_OPTIONS.  Replaces the reference's
// src/gpu/pdhg_solver.cu:166-411: the model is handed over as plain arrays
// (pm_lp), the solve runs on the device, and the SolveResult is filled
// exactly as the reference fills it (x and objective always, iterations =
// iter+1 when optimal else max_iterations, pv/dv of the final point,
// row_activity = A x in FP64).
#include <stdexcept>
#include <string>
#include <vector>

#include "abi_bridge.hpp"
#include "pramaan/gpu/pdhg_solver.hpp"
#include "pramaan_hip.h"

namespace pramaan {
namespace hip_abi {

SolveResult solve(const ModelIR& model, const pm_pdhg_options& o, std::vector<double>* y_out, double* kkt_out) {
    const int m = static_cast<int>(model.numRows());
    const int n = static_cast<int>(model.numVars());
    pm_lp lp{};
    lp.m = m;
    lp.n = n;
    lp.nnz = static_cast<int32_t>(model.A.nnz());
    lp.sense = model.obj_sense == ObjSense::kMaximize ? 1 : 0;
    lp.offset = model.obj_offset;
    lp.c = model.obj_coeffs.data();
    lp.row_ptr = model.A.rowPtr().data();
    lp.col_idx = model.A.colIdx().data();
    lp.values = model.A.values().data();
    lp.row_lower = model.row_lower.data();
    lp.row_upper = model.row_upper.data();
    lp.var_lower = model.var_lower.data();
    lp.var_upper = model.var_upper.data();
    const std::vector<CSRMatrix::Index> empty_rp(1, 0);
    if (model.A.rowPtr().empty()) lp.row_ptr = empty_rp.data();

    SolveResult result;
    result.x.assign(static_cast<size_t>(n), 0.0);
    result.row_activity.assign(static_cast<size_t>(m), 0.0);
    std::vector<double> y(static_cast<size_t>(m), 0.0);
    pm_pdhg_result r{};
    r.x = n > 0 ? result.x.data() : nullptr;
    r.y = m > 0 ? y.data() : nullptr;
    r.row_activity = m > 0 ? result.row_activity.data() : nullptr;
    const int rc = pm_pdhg_solve(&lp, &o, &r);
    if (rc != PM_OK) throw std::runtime_error(std::string("HIP PDHG error: ") + pm_last_error());
    switch (r.status) {
        case PM_SOLVE_OPTIMAL: result.status = SolveStatus::kOptimal; break;
        case PM_SOLVE_INFEASIBLE: result.status = SolveStatus::kInfeasible; break;
        case PM_SOLVE_UNBOUNDED: result.status = SolveStatus::kUnbounded; break;
        case PM_SOLVE_NUMERICAL_FAILURE: result.status = SolveStatus::kNumericalFailure; break;
        default: result.status = SolveStatus::kIterationLimit; break;
    }
    if (n == 0) result.x.clear();
    if (result.status == SolveStatus::kNumericalFailure) {
        // The reference returns before touching x (pdhg_solver.cu:186-192).
        result.x.clear();
        result.row_activity.clear();
    }
    result.iterations = r.iterations;
    result.objective_value = r.objective;
    result.primal_residual = r.primal_residual;
    result.dual_residual = r.dual_residual;
    if (y_out) *y_out = std::move(y);
    if (kkt_out) *kkt_out = r.rel_kkt;
    return result;
}

}  // namespace hip_abi

namespace gpu {

PdhgSolver::PdhgSolver(const PdhgOptions& opts) : options_(opts) {}
PdhgSolver::~PdhgSolver() {}

SolveResult PdhgSolver::solve(const ModelIR& model) {
    pm_pdhg_options o;
    pm_pdhg_options_default(&o);
    o.tolerance = options_.tolerance;
    o.max_iterations = options_.max_iterations;
    o.check_frequency = options_.check_frequency;
#ifdef PRAMAAN_HIP_EXTENDED_OPTIONS
    const int m = static_cast<int>(model.numRows());
    const int n = static_cast<int>(model.numVars());
    if (options_.pdlp) o.algorithm = PM_ALG_PDLP;
    if (options_.relative_kkt) o.termination = PM_TERM_RELATIVE_KKT;
    o.spmv_variant = options_.spmv_variant;
    o.max_refinements = options_.max_refinements;
    o.feasibility_polishing = options_.feasibility_polishing ? 1 : 0;
    o.device = options_.device;
    if (static_cast<int>(options_.x0.size()) == n && n > 0) o.x0 = options_.x0.data();
    if (static_cast<int>(options_.y0.size()) == m && m > 0) o.y0 = options_.y0.data();
    return hip_abi::solve(model, o, &last_y_, &last_kkt_);
#else
    return hip_abi::solve(model, o, nullptr, nullptr);
#endif
}

}  // namespace gpu
}  // namespace pramaan
