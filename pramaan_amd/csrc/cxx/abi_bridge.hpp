// abi_bridge.hpp -- one ModelIR solved through pm_pdhg_solve (internal to
// the C++ layer: gpu::PdhgSolver and make_hip_backend() both use it).
#pragma once

#include <vector>

#include "pramaan/ir.hpp"
#include "pramaan/simplex.hpp"
#include "pramaan_hip.h"

namespace pramaan {
namespace hip_abi {

// Solves `model` with options `o` on the device and fills a SolveResult the
// way the reference's PdhgSolver::solve does (src/gpu/pdhg_solver.cu:385-410).
// y_out / kkt_out (may be null) receive the dual and the relative KKT error.
// Engine errors become std::runtime_error.
SolveResult solve(const ModelIR& model, const pm_pdhg_options& o, std::vector<double>* y_out, double* kkt_out);

}  // namespace hip_abi
}  // namespace pramaan
