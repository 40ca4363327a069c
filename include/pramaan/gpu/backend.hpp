// backend.hpp -- compute-backend interface of PRAMAAN with the HIP backend.
//
// Backend is the reference's abstract interface
// (include/pramaan/gpu/backend.hpp:36-61).  make_hip_backend() is the
// implementation its comment names as the roadmap "HipBackend": it maps
// RevisedSimplex::Options to the GPU PDHG (max_iterations -> iteration
// limit, tolerance -> relative KKT tolerance of the FP64-refined solve) and
// creates a device context per call, so concurrent calls on one instance
// with different models are safe.
#pragma once

#include <memory>

#include "pramaan/ir.hpp"
#include "pramaan/simplex.hpp"

namespace pramaan {

class Backend {
public:
    virtual ~Backend() = default;
    virtual SolveResult solve_lp_relaxation(const ModelIR& model, const RevisedSimplex::Options& options) = 0;
};

}  // namespace pramaan

#include "pramaan/gpu/hip_backend.hpp"  // make_hip_backend()
