// hip_backend.hpp -- make_hip_backend(): the reference's Backend interface
// (include/pramaan/gpu/backend.hpp:36-61) on the GPU PDHG engine, the
// "HipBackend" its roadmap comment names.  The reference header declares
// only make_cpu_backend(), so callers built against it include this header
// for the HIP factory; it needs nothing beyond the reference's Backend,
// ModelIR and RevisedSimplex::Options.
//
// Mapping: RevisedSimplex::Options::max_iterations (a pivot cap) -> a PDHG
// budget of 100 iterations per pivot, at most 10^7; tolerance -> relative
// KKT target of the FP64-refined PDHG (at least 1e-10);
// primal_start_hint -> warm start.  Every call builds its own device
// context and holds no state in the instance, so concurrent calls on one
// instance are safe (the baseline the interface asks for).
#pragma once

#include <memory>

#include "pramaan/gpu/backend.hpp"

namespace pramaan {

std::unique_ptr<Backend> make_hip_backend();

}  // namespace pramaan
