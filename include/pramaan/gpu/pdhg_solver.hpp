// pdhg_solver.hpp -- the GPU PDHG backend of the PRAMAAN interface.
//
// Drop-in for the reference's gpu::PdhgSolver
// (include/pramaan/gpu/pdhg_solver.hpp:9-24): same class, constructor and
// non-const solve(const ModelIR&).  PdhgOptions keeps the reference's three
// fields and defaults; the fields after them select the engine's extensions
// and default to the reference behaviour.  The implementation
// (pramaan_amd/csrc/cxx/pdhg_solver.cpp) calls the C-ABI of
// include/pramaan_hip.h; HIP errors become std::runtime_error, as CUDA
// errors do in the reference (src/gpu/cuda_utils.cuh:8-15).
#pragma once

#include <vector>

#include "pramaan/ir.hpp"
#include "pramaan/simplex.hpp"

namespace pramaan {
namespace gpu {

// Marks the extended option/solver layout of this header.
#define PRAMAAN_HIP_EXTENDED_OPTIONS 1

struct PdhgOptions {
    double tolerance = 1e-6;
    int max_iterations = 100000;
    int check_frequency = 100;
    // ---- extensions (defaults = reference algorithm) ----
    bool pdlp = false;               // reflected Halpern PDHG + FP64 refinement
    bool relative_kkt = false;       // terminate on the FP64 relative KKT error
    int spmv_variant = 0;            // PM_SPMV_* (0 = autotuned)
    int max_refinements = 8;
    bool feasibility_polishing = true;  // PDLP: polish primal and dual feasibility once the gap is small
    int device = 0;
    std::vector<double> x0, y0;      // optional warm start (y: reduced cost c + A^T y)
};

class PdhgSolver {
public:
    PdhgSolver(const PdhgOptions& options = PdhgOptions());
    ~PdhgSolver();

    SolveResult solve(const ModelIR& model);

    // Dual solution and relative KKT error of the last solve (extensions).
    const std::vector<double>& lastDual() const { return last_y_; }
    double lastRelativeKkt() const { return last_kkt_; }

private:
    PdhgOptions options_;
    std::vector<double> last_y_;
    double last_kkt_ = 0.0;
};

}  // namespace gpu
}  // namespace pramaan
