// test_backend_hip.cpp -- make_hip_backend() (include/pramaan/gpu/hip_backend.hpp)
// through the reference's Backend interface (include/pramaan/gpu/backend.hpp
// of the reference, built by oracle/Makefile against its headers).  Checks the
// interface's stated baseline (backend.hpp:33-35: concurrent calls on one
// instance with different models are safe) the way the reference checks its
// CPU backend (tests/test_backend_cpu.cpp:367-397), plus known optima and the
// option mapping.  Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "pramaan/gpu/hip_backend.hpp"
#include "pramaan/ir.hpp"
#include "pramaan/simplex.hpp"

using namespace pramaan;

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
    std::printf("  %s %s\n", ok ? "ok:  " : "FAIL:", what.c_str());
    if (!ok) ++failures;
}

bool close_rel(double a, double b, double rel) { return std::fabs(a - b) <= rel * (1.0 + std::fabs(b)); }

// An LP from a dense row-major matrix.
ModelIR make_lp(ObjSense sense, std::vector<double> c, const std::vector<std::vector<double>>& rows,
                std::vector<double> rl, std::vector<double> ru, std::vector<double> xl, std::vector<double> xu) {
    const int n = static_cast<int>(c.size());
    std::vector<CSRMatrix::Index> rp{0}, ci;
    std::vector<double> v;
    for (const auto& r : rows) {
        for (int j = 0; j < n; ++j)
            if (r[j] != 0.0) {
                ci.push_back(j);
                v.push_back(r[j]);
            }
        rp.push_back(static_cast<CSRMatrix::Index>(ci.size()));
    }
    const std::size_t m = rows.size();
    return ModelIR(sense, 0.0, std::move(c), CSRMatrix(rp, ci, v, n), std::move(rl), std::move(ru),
                   std::vector<std::string>(m, "r"), std::move(xl), std::move(xu),
                   std::vector<VarType>(static_cast<std::size_t>(n), VarType::kContinuous),
                   std::vector<std::string>(static_cast<std::size_t>(n), "x"));
}

const double kInf = 1e30;

// max 3x + 5y s.t. x <= 4, 2y <= 12, 3x + 2y <= 18: optimum 36 at (2, 6).
ModelIR wyndor() {
    return make_lp(ObjSense::kMaximize, {3, 5}, {{1, 0}, {0, 2}, {3, 2}}, {-kInf, -kInf, -kInf}, {4, 12, 18},
                   {0, 0}, {kInf, kInf});
}
// min 2x + 3y s.t. x + y >= 1: optimum 2 at (1, 0).
ModelIR simple_min() { return make_lp(ObjSense::kMinimize, {2, 3}, {{1, 1}}, {1}, {kInf}, {0, 0}, {kInf, kInf}); }
// min -x - 2y s.t. x + y = 4, 1 <= x - y <= 2, 0 <= x, y <= 3: with y = 4 - x the
// objective is x - 8 and x - y = 2x - 4 in [1, 2], so the optimum is -5.5 at (2.5, 1.5).
ModelIR ranged() {
    return make_lp(ObjSense::kMinimize, {-1, -2}, {{1, 1}, {1, -1}}, {4, 1}, {4, 2}, {0, 0}, {3, 3});
}

}  // namespace

int main() {
    auto backend = make_hip_backend();
    expect(backend != nullptr, "make_hip_backend() returns an instance");
    RevisedSimplex::Options opts;  // reference defaults: 10,000 pivots, tolerance 1e-9

    const std::vector<std::pair<std::string, ModelIR>> models = {
        {"wyndor", wyndor()}, {"simple_min", simple_min()}, {"ranged", ranged()}};
    const double want[] = {36.0, 2.0, -5.5};
    std::vector<SolveResult> seq;
    for (std::size_t k = 0; k < models.size(); ++k) {
        seq.push_back(backend->solve_lp_relaxation(models[k].second, opts));
        expect(seq.back().status == SolveStatus::kOptimal, models[k].first + " optimal");
        expect(close_rel(seq.back().objective_value, want[k], 1e-7),
               models[k].first + " objective " + std::to_string(seq.back().objective_value));
    }

    // A pivot cap of 1 maps to a 100-iteration PDHG budget: not enough here.
    RevisedSimplex::Options tight;
    tight.max_iterations = 1;
    const SolveResult lim = backend->solve_lp_relaxation(wyndor(), tight);
    expect(lim.status == SolveStatus::kIterationLimit && lim.iterations == 100, "pivot cap 1 -> 100-iteration limit");
    // A very large pivot cap is clamped instead of overflowing.
    RevisedSimplex::Options huge;
    huge.max_iterations = 2000000000;
    expect(backend->solve_lp_relaxation(simple_min(), huge).status == SolveStatus::kOptimal, "huge pivot cap");
    // The primal start hint warm-starts the solve.
    RevisedSimplex::Options hinted;
    hinted.primal_start_hint = seq[0].x;
    const SolveResult warm = backend->solve_lp_relaxation(wyndor(), hinted);
    expect(warm.status == SolveStatus::kOptimal && warm.iterations <= seq[0].iterations, "start hint used");

    // Four threads share the instance, each with its own model (two run
    // the same one); every result equals the sequential one.
    const int kThreads = 4;
    std::vector<SolveResult> par(kThreads);
    std::vector<std::thread> pool;
    for (int t = 0; t < kThreads; ++t)
        pool.emplace_back([&, t] { par[t] = backend->solve_lp_relaxation(models[t % models.size()].second, opts); });
    for (auto& th : pool) th.join();
    for (int t = 0; t < kThreads; ++t) {
        const SolveResult& s = seq[t % models.size()];
        expect(par[t].status == SolveStatus::kOptimal && par[t].objective_value == s.objective_value &&
                   par[t].iterations == s.iterations && par[t].x == s.x,
               "thread " + std::to_string(t) + " equals the sequential solve of " + models[t % models.size()].first);
    }
    std::printf("%s\n", failures == 0 ? "ALL PASSED" : "FAILED");
    return failures == 0 ? 0 : 1;
}
