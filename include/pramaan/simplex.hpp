// simplex.hpp -- solve status and result types of the PRAMAAN interface.
//
// SolveStatus and SolveResult have the reference's members and order
// (include/pramaan/simplex.hpp:26-76), so every caller that switches on a
// status or reads x / objective_value / row_activity / residuals works with
// results of the GPU engine.  RevisedSimplex::Options is declared because the
// Backend interface (backend.hpp) takes it; the CPU simplex itself is not part
// of this engine.
#pragma once

#include <limits>
#include <vector>

#include "pramaan/ir.hpp"

namespace pramaan {

enum class SolveStatus {
    kOptimal,
    kInfeasible,
    kUnbounded,
    kIterationLimit,
    kNumericalFailure,
    kWarmStartRejected
};

struct SolveResult {
    SolveStatus status = SolveStatus::kIterationLimit;
    std::vector<double> x;
    double primal_residual = std::numeric_limits<double>::infinity();
    double dual_residual = std::numeric_limits<double>::infinity();
    double objective_value = 0.0;
    std::vector<double> row_activity;
    int iterations = 0;
    int crash_basis_pivots = 0;
    std::vector<bool> is_basic;
};

class RevisedSimplex {
public:
    struct Options {
        int max_iterations = 10000;
        double tolerance = 1e-9;
        std::vector<double> primal_start_hint;
    };
};

}  // namespace pramaan
