// gen_rolling.cpp -- rolling-horizon multi-period blending LPs (see the
// pm_rh_* declarations in pramaan_gen.h).
//
// One period (400 columns; a 200-row block plus 15 inventory-balance rows
// that link it to the period before, 215 rows in all):
//   columns  f[c][p]  c < 24 components, p < 15 products   (360, c-major)
//            s[p]     sales                                  (15)
//            I[p]     end-of-period product inventory        (15)
//            o[u]     overtime of processing unit u < 10     (10)
//   rows     quality  sum_c (a[c][q] - qmax[p][q]) f[c][p] <= 0   (15 x 8 = 120)
//            minimum  sum_c (a[c][q] - qmin[p][q]) f[c][p] >= 0   (7 x 3 = 21:
//                     products p < 7, attributes q < 3)
//            supply   sum_p f[c][p] <= avail[c][t]                 (24)
//            capacity sum_{p in unit u} sum_c f[c][p] - o[u] <= cap[u]  (10)
//            market   sum_{p in unit u} s[p] <= market[u]                 (10)
//            balance  I[p][t] - I[p][t-1] - sum_c f[c][p] + s[p] = 0    (15)
//                     (first period of a window: I[p][t-1] is the exogenous
//                      opening stock, moved to the right-hand side)
//            demand   s[p] <= d[p][t]                                     (15)
//   objective (min)  sum cost[c] f - price[p] s + hold[p] I + otcost[u] o
//   bounds   f, s >= 0; 0 <= I <= store[p]; 0 <= o <= otmax[u].
// Blend coefficients a ~ U(0.5, 1.5); demand d = base_p (1 + 0.4 sin(2 pi t
// / 168 + phase_p) + N(0, 0.1)), clipped at 0; qmax[p][q] >= 1.03 x the
// average blend + 0.01, and qmin[p][q] = the average blend times
// U(0.98, 1.02), both from the unperturbed coefficients, so
// about half the minimum rows bind where a product is made.  The zero plan
// (I = opening stock) satisfies every row: every window is feasible, and
// bounded because supply, demand and the bounds cap it.
// Perturbations (demand_noise, blend_noise) scale d and a per entry by
// U(1 - noise, 1 + noise) from their own random stream (perturb_seed); the
// rest of the plan depends on seed alone, so a perturbed plan is the
// unperturbed one with its demands and blend coefficients moved.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "pramaan_gen.h"

namespace {

constexpr int C = 24, P = 15, Q = 8, U = 10;
constexpr int PMIN = 7, QMIN = 3;                         // minimum-quality rows
constexpr int kCols = C * P + P + P + U;                  // 400
constexpr int kBlock = P * Q + PMIN * QMIN + C + U + U + P;  // 200
constexpr int kRows = kBlock + P;                         // 215 with the balance rows

struct Rng {  // xoshiro256** seeded by SplitMix64, as in gen.cpp
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            w = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * uniform());
    }
};

int col_f(int c, int p) { return c * P + p; }
int col_s(int p) { return C * P + p; }
int col_I(int p) { return C * P + P + p; }
int col_o(int u) { return C * P + 2 * P + u; }

}  // namespace

struct pm_rh {
    pm_rh_params prm;
    int T = 0;
    // per period
    std::vector<double> a;        // [T][C][Q] blend coefficients
    std::vector<double> qmax;     // [P][Q]
    std::vector<double> qmin;     // [PMIN][QMIN]
    std::vector<double> avail;    // [T][C]
    std::vector<double> demand;   // [T][P]
    std::vector<double> opening;  // [T][P] exogenous opening stock of a window starting at t
    double cost[C], price[P], hold[P], store[P], cap[U], otcost[U], otmax[U], market[U];
};

// The window LP reuses pm_gen_lp (defined in gen.cpp) through this builder.
struct pm_gen_lp;
extern "C" pm_gen_lp* pm__gen_lp_from_arrays(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci,
                                              const double* v, const double* c, const double* rl, const double* ru,
                                              const double* xl, const double* xu);

extern "C" {

int pm_rh_params_default(pm_rh_params* p) {
    if (!p) return PM_ERR_INVALID_ARG;
    std::memset(p, 0, sizeof(*p));
    p->periods_total = 523;  // 500 windows of 24 periods shifted by 1
    p->window = 24;
    p->demand_noise = 0.0;
    p->blend_noise = 0.0;
    p->seed = 1;
    p->perturb_seed = 0;
    return PM_OK;
}

int pm_rh_create(const pm_rh_params* prm, pm_rh** out) {
    if (!prm || !out) return PM_ERR_INVALID_ARG;
    if (prm->periods_total < 1 || prm->window < 1 || prm->window > prm->periods_total) return PM_ERR_INVALID_ARG;
    try {
        auto h = std::make_unique<pm_rh>();
        h->prm = *prm;
        const int T = prm->periods_total;
        h->T = T;
        Rng r(prm->seed);
        Rng pr(prm->perturb_seed ^ 0xB1E7D0ull);
        for (int c = 0; c < C; ++c) h->cost[c] = r.uniform(1.0, 3.0);
        for (int p = 0; p < P; ++p) {
            h->price[p] = r.uniform(4.0, 8.0);
            h->hold[p] = r.uniform(0.05, 0.2);
            h->store[p] = r.uniform(50.0, 150.0);
        }
        for (int u = 0; u < U; ++u) {
            h->cap[u] = r.uniform(200.0, 400.0);
            h->otcost[u] = r.uniform(0.5, 1.5);
            h->otmax[u] = 0.3 * h->cap[u];
        }
        std::vector<double> base_d(P), phase(P), base_a(C);
        for (int p = 0; p < P; ++p) { base_d[p] = r.uniform(80.0, 160.0); phase[p] = r.uniform(0.0, 6.283185307179586); }
        for (int u = 0; u < U; ++u) {
            double sd = 0.0;
            for (int p = u; p < P; p += U) sd += base_d[p];
            h->market[u] = 0.9 * sd;
        }
        for (int c = 0; c < C; ++c) base_a[c] = r.uniform(150.0, 300.0);
        h->a.resize(static_cast<size_t>(T) * C * Q);
        h->avail.resize(static_cast<size_t>(T) * C);
        h->demand.resize(static_cast<size_t>(T) * P);
        h->opening.resize(static_cast<size_t>(T) * P);
        std::vector<double> a0(C * Q);
        for (auto& v : a0) v = r.uniform(0.5, 1.5);
        h->qmax.resize(P * Q);
        for (int p = 0; p < P; ++p)
            for (int q = 0; q < Q; ++q) {
                double avg = 0.0;
                for (int c = 0; c < C; ++c) avg += a0[c * Q + q];
                avg /= C;
                h->qmax[p * Q + q] = std::max(r.uniform(1.0, 1.3), avg * 1.03 + 0.01);
            }
        // Minimum specifications come from a stream of their own, so every
        // other draw of the plan is the same with or without them.
        Rng rq(prm->seed ^ 0x9A11C0DEull);
        h->qmin.resize(PMIN * QMIN);
        for (int p = 0; p < PMIN; ++p)
            for (int q = 0; q < QMIN; ++q) {
                double avg = 0.0;
                for (int c = 0; c < C; ++c) avg += a0[c * Q + q];
                avg /= C;
                h->qmin[p * QMIN + q] = avg * rq.uniform(0.98, 1.02);
            }
        for (int t = 0; t < T; ++t) {
            const double season = 2.0 * 3.141592653589793 * t / 168.0;
            for (int c = 0; c < C; ++c) {
                h->avail[t * C + c] = base_a[c] * (1.0 + 0.2 * std::sin(season + 0.5 * c)) * r.uniform(0.9, 1.1);
                for (int q = 0; q < Q; ++q) {
                    double v = a0[c * Q + q] * r.uniform(0.97, 1.03);
                    if (prm->blend_noise > 0.0) v *= pr.uniform(1.0 - prm->blend_noise, 1.0 + prm->blend_noise);
                    h->a[(static_cast<size_t>(t) * C + c) * Q + q] = v;
                }
            }
            for (int p = 0; p < P; ++p) {
                double d = base_d[p] * (1.0 + 0.4 * std::sin(season + phase[p]) + 0.1 * r.normal());
                if (prm->demand_noise > 0.0) d *= pr.uniform(1.0 - prm->demand_noise, 1.0 + prm->demand_noise);
                h->demand[t * P + p] = d > 0.0 ? d : 0.0;
                h->opening[t * P + p] = h->store[p] * (0.3 + 0.2 * std::sin(season + phase[p] + 1.0));
            }
        }
        *out = h.release();
        return PM_OK;
    } catch (...) {
        return PM_ERR_OUT_OF_MEMORY;
    }
}

int pm_rh_window_size(const pm_rh* h, int32_t* rows_per_period, int32_t* cols_per_period, int32_t* windows) {
    if (!h) return PM_ERR_INVALID_ARG;
    if (rows_per_period) *rows_per_period = kRows;
    if (cols_per_period) *cols_per_period = kCols;
    if (windows) *windows = h->T - h->prm.window + 1;
    return PM_OK;
}

int pm_rh_window(const pm_rh* h, int32_t start, pm_gen_lp** out) {
    if (!h || !out) return PM_ERR_INVALID_ARG;
    const int W = h->prm.window;
    if (start < 0 || start + W > h->T) return PM_ERR_INVALID_ARG;
    const int n = W * kCols, m = W * kRows;
    std::vector<double> c(n), xl(n, 0.0), xu(n, PM_INFINITY_VALUE), rl(m), ru(m);
    std::vector<int32_t> rp(1, 0), ci;
    std::vector<double> v;
    auto row_end = [&](double lo, double hi, int i) {
        rl[i] = lo;
        ru[i] = hi;
        rp.push_back(static_cast<int32_t>(ci.size()));
    };
    auto put = [&](int col, double val) {
        ci.push_back(col);
        v.push_back(val);
    };
    for (int tau = 0; tau < W; ++tau) {
        const int t = start + tau, cb = tau * kCols, rb = tau * kRows;
        for (int cc = 0; cc < C; ++cc)
            for (int p = 0; p < P; ++p) c[cb + col_f(cc, p)] = h->cost[cc];
        for (int p = 0; p < P; ++p) {
            c[cb + col_s(p)] = -h->price[p];
            c[cb + col_I(p)] = h->hold[p];
            xu[cb + col_I(p)] = h->store[p];
        }
        for (int u = 0; u < U; ++u) {
            c[cb + col_o(u)] = h->otcost[u];
            xu[cb + col_o(u)] = h->otmax[u];
        }
        int i = rb;
        // quality rows (columns ascending: f[c][p] for c = 0..C-1)
        for (int p = 0; p < P; ++p)
            for (int q = 0; q < Q; ++q) {
                for (int cc = 0; cc < C; ++cc)
                    put(cb + col_f(cc, p), h->a[(static_cast<size_t>(t) * C + cc) * Q + q] - h->qmax[p * Q + q]);
                row_end(-PM_INFINITY_VALUE, 0.0, i++);
            }
        for (int p = 0; p < PMIN; ++p)  // minimum quality
            for (int q = 0; q < QMIN; ++q) {
                for (int cc = 0; cc < C; ++cc)
                    put(cb + col_f(cc, p), h->a[(static_cast<size_t>(t) * C + cc) * Q + q] - h->qmin[p * QMIN + q]);
                row_end(0.0, PM_INFINITY_VALUE, i++);
            }
        for (int cc = 0; cc < C; ++cc) {  // supply
            for (int p = 0; p < P; ++p) put(cb + col_f(cc, p), 1.0);
            row_end(-PM_INFINITY_VALUE, h->avail[t * C + cc], i++);
        }
        for (int u = 0; u < U; ++u) {  // capacity
            for (int cc = 0; cc < C; ++cc)
                for (int p = u; p < P; p += U) put(cb + col_f(cc, p), 1.0);
            put(cb + col_o(u), -1.0);
            row_end(-PM_INFINITY_VALUE, h->cap[u], i++);
        }
        for (int u = 0; u < U; ++u) {  // market
            for (int p = u; p < P; p += U) put(cb + col_s(p), 1.0);
            row_end(-PM_INFINITY_VALUE, h->market[u], i++);
        }
        for (int p = 0; p < P; ++p) {  // balance: [I(t-1)] f... s I(t)
            if (tau > 0) put(cb - kCols + col_I(p), -1.0);
            for (int cc = 0; cc < C; ++cc) put(cb + col_f(cc, p), -1.0);
            put(cb + col_s(p), 1.0);
            put(cb + col_I(p), 1.0);
            const double rhs = tau == 0 ? h->opening[static_cast<size_t>(t) * P + p] : 0.0;
            row_end(rhs, rhs, i++);
        }
        for (int p = 0; p < P; ++p) {  // demand
            put(cb + col_s(p), 1.0);
            row_end(-PM_INFINITY_VALUE, h->demand[t * P + p], i++);
        }
    }
    pm_gen_lp* g = pm__gen_lp_from_arrays(m, n, rp.data(), ci.data(), v.data(), c.data(), rl.data(), ru.data(),
                                          xl.data(), xu.data());
    if (!g) return PM_ERR_OUT_OF_MEMORY;
    *out = g;
    return PM_OK;
}

void pm_rh_free(pm_rh* h) { delete h; }

}  // extern "C"
