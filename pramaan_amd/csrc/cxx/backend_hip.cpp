// backend_hip.cpp -- make_hip_backend(): the Backend interface
// (reference include/pramaan/gpu/backend.hpp:36-61) on the GPU PDHG engine.
// Each call builds its own PdhgSolver (and device context), so concurrent
// calls on one Backend instance are independent.
#include "pramaan/gpu/backend.hpp"
#include "pramaan/gpu/pdhg_solver.hpp"

namespace pramaan {
namespace {

class HipBackend final : public Backend {
public:
    SolveResult solve_lp_relaxation(const ModelIR& model, const RevisedSimplex::Options& options) override {
        gpu::PdhgOptions o;
        o.pdlp = true;
        o.relative_kkt = true;
        o.check_frequency = 64;
        // The simplex pivot limit and its 1e-9 feasibility tolerance map to
        // a PDHG iteration budget and a relative KKT target.
        o.max_iterations = options.max_iterations > 0 ? options.max_iterations * 100 : 100000;
        o.tolerance = options.tolerance > 1e-10 ? options.tolerance : 1e-10;
        if (!options.primal_start_hint.empty()) o.x0 = options.primal_start_hint;
        gpu::PdhgSolver solver(o);
        return solver.solve(model);
    }
};

}  // namespace

std::unique_ptr<Backend> make_hip_backend() { return std::make_unique<HipBackend>(); }

}  // namespace pramaan
