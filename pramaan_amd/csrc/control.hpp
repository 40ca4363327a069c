// control.hpp -- the restart / refinement / termination logic of the
// reflected-Halpern PDHG, shared by the single-LP host loop (engine.hip)
// and the batched device kernel (batch.hip) so both make the same decision
// from the same FP64 KKT numbers.
//
// Every check_frequency iterations the caller evaluates the FP64 relative
// KKT error `rel` of the candidate T(z) (as a point of the original LP) and
// calls on_check(), which returns one action:
//   OPTIMAL  rel <= tol
//   LIMIT    iteration budget spent
//   FAIL     rel is not finite
//   REFINE   new iterative-refinement round around the candidate: the round
//            gained round_gain (1e-3) since its start, or it gained 2x and
//            20 checks in a row failed to halve the best error seen before
//            them (the FP32 floor, or a primal weight that ran away: a new
//            round re-centres the problem on the candidate)
//   RESTART  z <- T(z), anchor <- z: the error fell to 0.2x the error at the
//            last restart, or below 0.8x of it while rising again, or 36 % of
//            all iterations passed since (PDLP's adaptive restart rule), or,
//            once per round, the stall above without the gain a new round
//            needs (the restart updates the primal weight: a rolling-horizon
//            window stalled 5,000 iterations at 2.2e-6 with a dual residual
//            that only a 25x smaller weight, set at the next artificial
//            restart, unblocked)
//   CONTINUE otherwise.
// After RESTART the caller measures ||dx||, ||dy|| moved since the previous
// restart and calls on_restart() (primal weight update, theta = 0.5, kept
// within a factor kOmegaSpan = 100 of the initial weight: when x stops
// moving the update alone drives the weight up without bound -- 160 to
// 4.6e5 on a generated 3709 x 9273 LP whose gap then stalled at 1.0 -- and
// the bound solved it in 3k iterations with no cost on the others tried).
// A stall restart whose error is dominated by a feasibility residual
// (max(rel_p, rel_d) > rel_g) balances the residuals instead: omega (the
// dual step is eta * omega, the primal step eta / omega) moves by
// sqrt(rel_p / rel_d), at most kBalance either way.  In a stall dominated
// by the primal residual y barely moves, so the distance rule lowers omega
// and deepens the stall: a rolling-horizon window sat 14,000 iterations at
// rel_p 6.8e-6 (rel_d 1e-8) while the distance rule took omega from 7.5e-3
// to 4.4e-3, and left the stall in one check once a restart raised omega
// 18x.  A stall in the gap keeps the distance rule: on box-bounded LPs
// rel_d is zero by construction and the dual error shows only in the gap.
// After REFINE the caller calls nothing: on_check already set the new scale
// factors DP = DD = min(1/rel, 1e4 * previous) and reset the round state.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace pm {

struct Control {
    enum Action { CONTINUE = 0, OPTIMAL = 1, LIMIT = 2, FAIL = 3, REFINE = 4, RESTART = 5 };
    static constexpr int kStall = 20;
    static constexpr double kSuff = 0.2, kNec = 0.8, kArt = 0.36, kTheta = 0.5, kRoundGain = 1e-3,
                            kStallGain = 0.5, kOmegaSpan = 100.0, kBalance = 10.0;

    // parameters
    double tol;
    int max_iterations;
    int max_refinements;
    // state
    double DP, DD, omega, omega_init;
    int total_it, k, restarts, rounds, it_since;
    double k_round_start, k_last, k_prev;
    double ring[kStall];
    int nhist;
    double older_min;
    bool stall_restarted;  // the stall restart of this round was taken
    bool balance_next;     // the pending restart is a stall restart
    double last_p, last_d, last_g;  // relative primal, dual and gap errors of the last check
    bool use_g;

    __host__ __device__ void init(double tol_, int max_it, int max_ref, double omega0, bool warm_dual) {
        tol = tol_;
        max_iterations = max_it;
        max_refinements = max_ref;
        DP = DD = 1.0;
        omega = omega_init = omega0;
        total_it = k = restarts = rounds = it_since = 0;
        balance_next = false;
        last_p = last_d = last_g = 0.0;
        use_g = warm_dual;
        new_round();
    }

    __host__ __device__ void new_round() {
        k = 0;
        k_round_start = -1.0;
        k_last = -1.0;
        k_prev = INFINITY;
        it_since = 0;
        nhist = 0;
        older_min = INFINITY;
        stall_restarted = false;
    }

    // Iterations to run before the next check.
    __host__ __device__ int chunk(int check_frequency) const {
        const int left = max_iterations - total_it;
        return left < check_frequency ? left : check_frequency;
    }

    __host__ __device__ void advance(int iters) {
        k += iters;
        total_it += iters;
        it_since += iters;
    }

    // rel = max(rel_p, rel_d, rel_g) of the candidate.
    __host__ __device__ Action on_check(double rel, double rel_p, double rel_d, double rel_g) {
        if (!(rel == rel) || rel == INFINITY) return FAIL;
        last_p = rel_p;
        last_d = rel_d;
        last_g = rel_g;
        if (rel <= tol) return OPTIMAL;
        if (total_it >= max_iterations) return LIMIT;
        if (k_round_start < 0.0) k_round_start = rel;
        // stall window: ring of the last kStall errors, older_min of the rest
        bool stalled = false;
        if (nhist >= kStall) {
            const double out = ring[nhist % kStall];
            older_min = out < older_min ? out : older_min;
        }
        ring[nhist % kStall] = rel;
        ++nhist;
        if (nhist > kStall) {
            double recent = INFINITY;
            for (int q = 0; q < kStall; ++q) recent = ring[q] < recent ? ring[q] : recent;
            stalled = recent > 0.5 * older_min;
        }
        // A stall counts once the round has at least halved its error.
        if (rounds < max_refinements &&
            (rel <= k_round_start * kRoundGain || (stalled && rel <= k_round_start * kStallGain))) {
            double f = 1.0 / (rel > 1e-300 ? rel : 1e-300);
            const double cap = (DP > 1.0 ? DP : 1.0) * 1e4;
            f = f < cap ? f : cap;
            DP = DD = f > 1.0 ? f : 1.0;
            use_g = true;
            ++rounds;
            new_round();
            return REFINE;
        }
        if (k_last < 0.0) k_last = rel;
        const bool stall_restart = stalled && !stall_restarted;
        if (stall_restart) stall_restarted = true;
        const bool restart = rel <= kSuff * k_last || (rel <= kNec * k_last && rel > k_prev) ||
                             static_cast<double>(it_since) >= kArt * static_cast<double>(total_it) ||
                             stall_restart;
        k_prev = rel;
        if (restart) {
            k = 0;
            k_last = rel;
            k_prev = INFINITY;
            it_since = 0;
            ++restarts;
            balance_next = stall_restart;
            return RESTART;
        }
        return CONTINUE;
    }

    __host__ __device__ void on_restart(double dx, double dy) {
        const double big = last_p > last_d ? last_p : last_d;
        const bool balance = balance_next && big > last_g;
        balance_next = false;
        if (balance) {
            // Residuals are floored at 1e-3 of the larger, so the ratio's
            // square root stays finite before the kBalance cap.
            const double fl = 1e-3 * big;
            double f = sqrt((last_p > fl ? last_p : fl) / (last_d > fl ? last_d : fl));
            f = f < 1.0 / kBalance ? 1.0 / kBalance : (f > kBalance ? kBalance : f);
            omega *= f;
        } else if (dx > 1e-10 && dy > 1e-10) {
            omega = exp(kTheta * log(dy / dx) + (1.0 - kTheta) * log(omega));
        }
        const double lo = omega_init / kOmegaSpan, hi = omega_init * kOmegaSpan;
        omega = omega < lo ? lo : (omega > hi ? hi : omega);
    }

    // Halpern weight of the iteration with in-round index kk.
    __host__ __device__ static float lambda(int kk) {
        return static_cast<float>((kk + 1.0) / (kk + 2.0));
    }
};

}  // namespace pm
