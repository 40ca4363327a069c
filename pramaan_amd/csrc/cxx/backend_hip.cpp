// backend_hip.cpp -- make_hip_backend() (include/pramaan/gpu/hip_backend.hpp)
// over the engine's C-ABI.  It fills pm_pdhg_options directly, so it
// compiles against this project's headers and, unchanged, against the
// reference's (whose PdhgOptions has only its three fields).
#include <algorithm>
#include <cstdint>

#include "abi_bridge.hpp"
#include "pramaan/gpu/hip_backend.hpp"

namespace pramaan {
namespace {

constexpr std::int64_t kItersPerPivot = 100, kMaxIterations = 10000000;

class HipBackend final : public Backend {
public:
    SolveResult solve_lp_relaxation(const ModelIR& model, const RevisedSimplex::Options& options) override {
        pm_pdhg_options o;
        pm_pdhg_options_default(&o);
        o.algorithm = PM_ALG_PDLP;
        o.termination = PM_TERM_RELATIVE_KKT;
        o.check_frequency = 64;
        const std::int64_t budget =
            options.max_iterations > 0 ? static_cast<std::int64_t>(options.max_iterations) * kItersPerPivot : 100000;
        o.max_iterations = static_cast<int32_t>(std::min(budget, kMaxIterations));
        o.tolerance = std::max(options.tolerance, 1e-10);
        if (!options.primal_start_hint.empty() &&
            options.primal_start_hint.size() == static_cast<std::size_t>(model.numVars()))
            o.x0 = options.primal_start_hint.data();
        return hip_abi::solve(model, o, nullptr, nullptr);
    }
};

}  // namespace

std::unique_ptr<Backend> make_hip_backend() { return std::make_unique<HipBackend>(); }

}  // namespace pramaan
