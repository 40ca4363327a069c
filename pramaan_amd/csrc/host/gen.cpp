// gen.cpp -- synthetic LPs with a closed-form optimum (see pramaan_gen.h).
//
// Random numbers: xoshiro256** (Blackman & Vigna, 2018) seeded through
// SplitMix64, uniforms from the top 53 bits, normals by Box-Muller.  The
// draw order is fixed below, so a seed gives the same LP on any platform.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pramaan_gen.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s) {  // SplitMix64
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            w = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }  // bias < 2^-40 for our n
    double normal() {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace

struct pm_gen_lp {
    int32_t m = 0, n = 0;
    std::vector<int32_t> rp, ci;
    std::vector<double> v, c, rl, ru, xl, xu, xs, ys;
    double pstar = 0.0;
};

extern "C" {

int pm_gen_params_named(int32_t which, uint64_t seed, pm_gen_params* p) {
    if (!p) return PM_ERR_INVALID_ARG;
    std::memset(p, 0, sizeof(*p));
    p->support = 0.3;
    p->var_upper = 10.0;
    p->seed = seed;
    switch (which) {
        case 0:
            p->m = 2000; p->n = 5000; p->k_min = p->k_max = 8; p->eq_frac = 1.0; p->noise = 0.0;
            return PM_OK;
        case 1:
            p->m = 500000; p->n = 1000000; p->k_min = p->k_max = 12; p->eq_frac = 0.6; p->noise = 0.05;
            return PM_OK;
        case 2:
            p->m = 50000; p->n = 100000; p->k_min = p->k_max = 12; p->eq_frac = 0.6; p->noise = 0.05;
            return PM_OK;
        case 3: {
            Rng r(seed ^ 0x5A17BA7C4ull);
            const double lm = std::log(50.0) + (std::log(5000.0) - std::log(50.0)) * r.uniform();
            p->m = static_cast<int32_t>(std::lround(std::exp(lm)));
            p->n = static_cast<int32_t>(std::lround(2.5 * p->m));
            p->k_min = 4; p->k_max = 16; p->eq_frac = 1.0; p->noise = 0.0;
            return PM_OK;
        }
        default: return PM_ERR_INVALID_ARG;
    }
}

int pm_gen_synthetic(const pm_gen_params* p, pm_gen_lp** out) {
    if (!p || !out) return PM_ERR_INVALID_ARG;
    *out = nullptr;
    if (p->m < 0 || p->n < 0 || p->k_min < 0 || p->k_max < p->k_min) return PM_ERR_INVALID_ARG;
    try {
        auto g = std::make_unique<pm_gen_lp>();
        const int m = p->m, n = p->n;
        g->m = m;
        g->n = n;
        Rng r(p->seed);
        // 1. Columns: k_j distinct rows (Floyd's sampling), sorted; values.
        std::vector<int32_t> colptr(static_cast<size_t>(n) + 1, 0), rows;
        std::vector<double> vals;
        rows.reserve(static_cast<size_t>(n) * static_cast<size_t>(p->k_max));
        vals.reserve(rows.capacity());
        std::vector<int32_t> pick;
        for (int j = 0; j < n; ++j) {
            int k = p->k_min + static_cast<int>(r.below(static_cast<uint64_t>(p->k_max - p->k_min + 1)));
            if (k > m) k = m;
            pick.clear();
            for (int t = m - k; t < m; ++t) {  // Floyd: k distinct of [0, m)
                const int32_t cand = static_cast<int32_t>(r.below(static_cast<uint64_t>(t) + 1));
                if (std::find(pick.begin(), pick.end(), cand) == pick.end()) pick.push_back(cand);
                else pick.push_back(t);
            }
            std::sort(pick.begin(), pick.end());
            for (int32_t i : pick) {
                double a = r.normal();
                if (p->noise > 0.0) a *= 1.0 + r.uniform(-p->noise, p->noise);
                rows.push_back(i);
                vals.push_back(static_cast<double>(static_cast<float>(a)));
            }
            colptr[j + 1] = static_cast<int32_t>(rows.size());
        }
        const size_t nnz = rows.size();
        // 2. CSR by counting sort over columns in order (rows stay sorted).
        g->rp.assign(static_cast<size_t>(m) + 1, 0);
        for (size_t k = 0; k < nnz; ++k) ++g->rp[static_cast<size_t>(rows[k]) + 1];
        for (int i = 0; i < m; ++i) g->rp[i + 1] += g->rp[i];
        g->ci.resize(nnz);
        g->v.resize(nnz);
        {
            std::vector<int32_t> cur(g->rp.begin(), g->rp.end() - 1);
            for (int j = 0; j < n; ++j)
                for (int k = colptr[j]; k < colptr[j + 1]; ++k) {
                    const int d = cur[rows[k]]++;
                    g->ci[d] = j;
                    g->v[d] = vals[k];
                }
        }
        // 3. x*: partial Fisher-Yates picks the support.
        g->xs.assign(n, 0.0);
        {
            std::vector<int32_t> perm(n);
            for (int j = 0; j < n; ++j) perm[j] = j;
            const int ns = static_cast<int>(std::floor(p->support * n));
            for (int t = 0; t < ns; ++t) {
                const int u = t + static_cast<int>(r.below(static_cast<uint64_t>(n - t)));
                std::swap(perm[t], perm[u]);
                g->xs[perm[t]] = r.uniform();
            }
        }
        // 4. Rows: equality subset, slacks, duals.
        std::vector<char> is_eq(m, 1);
        {
            std::vector<int32_t> perm(m);
            for (int i = 0; i < m; ++i) perm[i] = i;
            const int neq = static_cast<int>(std::lround(p->eq_frac * m));
            for (int t = 0; t < m; ++t) {
                const int u = t + static_cast<int>(r.below(static_cast<uint64_t>(m - t)));
                std::swap(perm[t], perm[u]);
            }
            for (int t = neq; t < m; ++t) is_eq[perm[t]] = 0;
        }
        g->rl.resize(m);
        g->ru.resize(m);
        g->ys.assign(m, 0.0);
        for (int i = 0; i < m; ++i) {
            double ax = 0.0;
            for (int k = g->rp[i]; k < g->rp[i + 1]; ++k) ax += g->v[k] * g->xs[g->ci[k]];
            if (is_eq[i]) {
                g->rl[i] = g->ru[i] = ax;
                g->ys[i] = r.normal();
            } else {
                g->rl[i] = -PM_INFINITY_VALUE;
                g->ru[i] = ax + r.uniform();
            }
        }
        // 5. c = A^T y* + z.
        g->c.assign(n, 0.0);
        for (int i = 0; i < m; ++i)
            for (int k = g->rp[i]; k < g->rp[i + 1]; ++k) g->c[g->ci[k]] += g->v[k] * g->ys[i];
        for (int j = 0; j < n; ++j) {
            const double z = g->xs[j] > 0.0 ? 0.0 : r.uniform();
            g->c[j] += z;
        }
        g->xl.assign(n, 0.0);
        g->xu.assign(n, p->var_upper);
        double ps = 0.0;
        for (int j = 0; j < n; ++j) ps += g->c[j] * g->xs[j];
        g->pstar = ps;
        *out = g.release();
        return PM_OK;
    } catch (const std::bad_alloc&) {
        return PM_ERR_OUT_OF_MEMORY;
    } catch (...) {
        return PM_ERR_INTERNAL;
    }
}

int pm_gen_view(const pm_gen_lp* g, pm_lp* lp, const double** x_star, const double** y_star, double* p_star) {
    if (!g || !lp) return PM_ERR_INVALID_ARG;
    lp->m = g->m;
    lp->n = g->n;
    lp->nnz = static_cast<int32_t>(g->v.size());
    lp->sense = 0;
    lp->offset = 0.0;
    lp->c = g->c.data();
    lp->row_ptr = g->rp.data();
    lp->col_idx = g->ci.data();
    lp->values = g->v.data();
    lp->row_lower = g->rl.data();
    lp->row_upper = g->ru.data();
    lp->var_lower = g->xl.data();
    lp->var_upper = g->xu.data();
    if (x_star) *x_star = g->xs.data();
    if (y_star) *y_star = g->ys.data();
    if (p_star) *p_star = g->pstar;
    return PM_OK;
}

void pm_gen_free(pm_gen_lp* g) { delete g; }

// Internal: wraps LP arrays (no known optimum) in a pm_gen_lp.
pm_gen_lp* pm__gen_lp_from_arrays(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v,
                                  const double* c, const double* rl, const double* ru, const double* xl,
                                  const double* xu) {
    try {
        auto g = std::make_unique<pm_gen_lp>();
        g->m = m;
        g->n = n;
        const int nnz = rp[m];
        g->rp.assign(rp, rp + m + 1);
        g->ci.assign(ci, ci + nnz);
        g->v.assign(v, v + nnz);
        g->c.assign(c, c + n);
        g->rl.assign(rl, rl + m);
        g->ru.assign(ru, ru + m);
        g->xl.assign(xl, xl + n);
        g->xu.assign(xu, xu + n);
        g->pstar = NAN;
        return g.release();
    } catch (...) {
        return nullptr;
    }
}

}  // extern "C"
