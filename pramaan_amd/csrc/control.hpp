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
#include <cstdint>

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

    // rel = max(rel_p, rel_d, rel_g) of the candidate.  q >= 0: the quantity
    // the restart rule compares instead of rel (the normalized duality gap,
    // for LPs whose every column is boxed); termination, refinement rounds
    // and stalls always use rel.
    __host__ __device__ Action on_check(double rel, double rel_p, double rel_d, double rel_g, double q = -1.0) {
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
        const double r = q >= 0.0 ? q : rel;
        if (k_last < 0.0) k_last = r;
        const bool stall_restart = stalled && !stall_restarted;
        if (stall_restart) stall_restarted = true;
        const bool restart = r <= kSuff * k_last || (r <= kNec * k_last && r > k_prev) ||
                             static_cast<double>(it_since) >= kArt * static_cast<double>(total_it) ||
                             stall_restart;
        k_prev = r;
        if (restart) {
            k = 0;
            k_last = r;
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

// Feasibility polishing (after PDLP's): once a check finds the relative gap
// small (<= kGap tol) and the error within kRel tol (at most once every
// kEvery iterations), the primal and the dual feasibility problems are
// solved on their own from the candidate,
// each by the same reflected-Halpern PDHG with the main primal weight:
//   PRIMAL  the LP with cost 0, from x_cand (its dual starts at 0): stops
//           once the primal residual of x_tot is <= kTarget tol;
//   DUAL    the LP with every finite bound moved to 0, from y_cand (x at
//           0): stops once the dual residual of y_tot is <= kTarget tol.
// Each phase restarts on its own residual by PDLP's rule and has a budget
// of kFrac of the iterations so far (at least kMinBudget).  When both meet
// their targets the pair (x_primal, y_dual) is accepted if its relative KKT
// error meets the tolerance and its relative gap is at most kAccept tol:
// both points are then feasible to 0.3 tol, the optimum lies between their
// objectives, and a gap of half the tolerance keeps the primal objective
// within about tol (1 + |objective|) of it -- the KKT test alone would let
// it drift to twice that (a polished rolling-horizon window ended 1.01e-6
// from the HiGHS objective).  Otherwise, or when a phase runs out of
// budget, the main solve continues exactly where it was when the polish
// began (its iterates, round data and control state are held meanwhile),
// so a polish that fails costs only its own iterations.  A phase that
// starts from a residual already at its target is skipped.  Polishing
// targets the solves whose gap has converged while a feasibility residual
// stalls; the thresholds come from a numpy restatement of this iteration
// on 20 rolling-horizon windows of 5,160 x 9,600 (mean iterations 5,261
// without polishing, 4,067 with these; accepting on the KKT error alone
// would give 3,398 at the objective accuracy cost above).
struct Polish {
    enum Phase { MAIN = 0, PRIMAL = 1, DUAL = 2, VERIFY = 3 };  // VERIFY: the pair's KKT is being evaluated
    enum Action { CONTINUE = 0, DONE = 1, FAILED = 2, RESTART = 3 };
    static constexpr double kGap = 0.1, kRel = 3.0, kTarget = 0.3, kFrac = 0.1, kAccept = 0.5;
    static constexpr int kEvery = 512, kMinBudget = 256;

    int32_t enabled, phase, k, it, since, budget, last_try, attempts, fails;
    double q_last, q_prev;
    double save_rel, save_p, save_d;  // errors of the saved candidate

    __host__ __device__ void init(bool on) {
        enabled = on ? 1 : 0;
        phase = MAIN;
        k = it = since = budget = attempts = fails = 0;
        last_try = -(1 << 30);
        q_last = -1.0;
        q_prev = INFINITY;
        save_rel = save_p = save_d = 0.0;
    }
    // Whether a main-solve check of (rel, rel_g) at total_it starts polishing;
    // each failed attempt doubles the wait before the next (up to 64x).
    __host__ __device__ bool trigger(double tol, int total_it, double rel, double rel_g) const {
        const int wait = kEvery << (fails < 6 ? fails : 6);
        return enabled && phase == MAIN && rel > tol && rel <= kRel * tol && rel_g <= kGap * tol &&
               total_it - last_try >= wait;
    }
    __host__ __device__ void save(int total_it, double rel, double rel_p, double rel_d) {
        ++attempts;
        last_try = total_it;
        save_rel = rel;
        save_p = rel_p;
        save_d = rel_d;
    }
    __host__ __device__ bool needs(double tol, double q) const { return q > kTarget * tol; }
    // Whether a polished pair of relative KKT error rel and gap rel_g ends the solve.
    __host__ __device__ static bool accept(double tol, double rel, double rel_g) {
        return rel <= tol && rel_g <= kAccept * tol;
    }
    __host__ __device__ void start(int ph, int total_it) {
        phase = ph;
        k = it = since = 0;
        const int b = static_cast<int>(kFrac * total_it);
        budget = b > kMinBudget ? b : kMinBudget;
        q_last = -1.0;
        q_prev = INFINITY;
    }
    // Check of a phase after `iters` more iterations, q its residual.
    __host__ __device__ Action on_check(double tol, double q, int iters) {
        k += iters;
        it += iters;
        since += iters;
        if (q <= kTarget * tol) return DONE;
        if (!(q == q) || it >= budget) return FAILED;
        if (q_last < 0.0) q_last = q;
        const bool restart = q <= Control::kSuff * q_last || (q <= Control::kNec * q_last && q > q_prev) ||
                             static_cast<double>(since) >= Control::kArt * static_cast<double>(it);
        q_prev = q;
        if (restart) {
            k = 0;
            since = 0;
            q_last = q;
            q_prev = INFINITY;
            return RESTART;
        }
        return CONTINUE;
    }
    // Scale factor of a phase's correction problem that starts from residual q.
    __host__ __device__ static double scale(double q, double dp) {
        double f = 1.0 / (q > 1e-300 ? q : 1e-300);
        const double cap = (dp > 1.0 ? dp : 1.0) * 1e4;
        f = f < cap ? f : cap;
        return f > 1.0 ? f : 1.0;
    }
};

}  // namespace pm
