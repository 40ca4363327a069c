// test_host_sanitize.cpp -- this project's host-side C and C++ code under
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/cxx/Makefile,
// target `sanitize`; no GPU code is linked).  It drives:
//   * the synthetic generator (pm_gen_*) over its named workloads and
//     custom sizes, the empty and one-column edge cases included;
//   * the rolling-horizon generator (pm_rh_*): nominal and perturbed
//     plans, first, middle and last windows, a window past the end
//     (rejected);
//   * the network's host feature builder (pm_gnn_lp_features) on windows
//     and generated LPs, with and without period lengths;
//   * the certificate writer and checker (pm_cert_write / pm_cert_verify):
//     V1 and V2 of a generated LP's known optimum accepted, a tampered x
//     rejected, an unreadable file reported.
// Any sanitizer report aborts the process (halt_on_error); the program
// prints ALL PASSED only when every check holds.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "pramaan_cert.h"
#include "pramaan_gen.h"
#include "pramaan_gnn.h"
#include "pramaan_hip.h"

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
    if (!ok) {
        ++failures;
        std::printf("FAILED: %s\n", what.c_str());
    }
}

void features(const pm_lp& lp, int pc, int pr, const std::string& tag) {
    const size_t nz = lp.nnz > 0 ? lp.nnz : 1;
    std::vector<int32_t> cp(lp.n + 1), ri(nz);
    std::vector<float> ar(nz), ac(nz), xr(static_cast<size_t>(lp.m > 0 ? lp.m : 1) * 16),
        xc(static_cast<size_t>(lp.n) * 16);
    double sc[2] = {0.0, 0.0};
    const int rc = pm_gnn_lp_features(&lp, pc, pr, cp.data(), ri.data(), ar.data(), ac.data(), xr.data(),
                                      xc.data(), sc);
    expect(rc == 0, tag + ": features built");
    expect(cp[lp.n] == lp.nnz, tag + ": column pointers end at nnz");
    bool finite = std::isfinite(sc[0]) && std::isfinite(sc[1]) && sc[0] > 0.0 && sc[1] > 0.0;
    for (float v : xc) finite = finite && std::isfinite(v);
    expect(finite, tag + ": finite features and scales");
}

}  // namespace

int main() {
    // synthetic LPs
    for (int which : {0, 2, 3}) {
        for (uint64_t seed : {1ull, 7ull}) {
            pm_gen_params p;
            expect(pm_gen_params_named(which, seed, &p) == 0, "named parameters");
            pm_gen_lp* g = nullptr;
            expect(pm_gen_synthetic(&p, &g) == 0 && g != nullptr, "generate workload " + std::to_string(which));
            pm_lp lp;
            const double *xs = nullptr, *ys = nullptr;
            double ps = 0.0;
            expect(pm_gen_view(g, &lp, &xs, &ys, &ps) == 0, "view");
            double obj = 0.0;
            for (int j = 0; j < lp.n; ++j) obj += lp.c[j] * xs[j];
            expect(std::fabs(obj - ps) <= 1e-9 * (1.0 + std::fabs(ps)), "closed-form objective");
            features(lp, 0, 0, "generated " + std::to_string(which));
            if (which == 0 && seed == 1) {
                // certificates of the known optimum (y in the reference sign: -y*)
                std::vector<double> yref(lp.m), xbad(xs, xs + lp.n);
                for (int i = 0; i < lp.m; ++i) yref[i] = -ys[i];
                char err[512], msg[4096];
                const std::string v1 = "host_sanitize_v1.cert", v2 = "host_sanitize_v2.cert",
                                  bad = "host_sanitize_bad.cert";
                expect(pm_cert_write(&lp, xs, nullptr, "OPTIMAL", 1e-6, v1.c_str(), err, sizeof err) == 0, "write V1");
                expect(pm_cert_write(&lp, xs, yref.data(), "OPTIMAL", 1e-6, v2.c_str(), err, sizeof err) == 0,
                       "write V2");
                expect(pm_cert_verify(&lp, v1.c_str(), msg, sizeof msg) == 0, std::string("V1 accepted: ") + msg);
                expect(pm_cert_verify(&lp, v2.c_str(), msg, sizeof msg) == 0, std::string("V2 accepted: ") + msg);
                xbad[0] += 1.0;
                expect(pm_cert_write(&lp, xbad.data(), yref.data(), "OPTIMAL", 1e-6, bad.c_str(), err, sizeof err) == 0,
                       "write tampered");
                expect(pm_cert_verify(&lp, bad.c_str(), msg, sizeof msg) == 1, "tampered x rejected");
                expect(pm_cert_verify(&lp, "no_such_file.cert", msg, sizeof msg) == 2, "missing file reported");
                std::remove(v1.c_str());
                std::remove(v2.c_str());
                std::remove(bad.c_str());
            }
            pm_gen_free(g);
        }
    }
    // edge cases: no rows, one column
    {
        pm_gen_params p;
        pm_gen_params_named(0, 3, &p);
        p.m = 0;
        p.n = 5;
        pm_gen_lp* g = nullptr;
        expect(pm_gen_synthetic(&p, &g) == 0, "zero-row LP generated");
        pm_lp lp;
        pm_gen_view(g, &lp, nullptr, nullptr, nullptr);
        expect(lp.m == 0 && lp.nnz == 0, "zero-row LP has no entries");
        pm_gen_free(g);
        p.m = 3;
        p.n = 1;
        p.k_min = p.k_max = 8;  // more than the rows: clamped
        expect(pm_gen_synthetic(&p, &g) == 0, "one-column LP generated");
        pm_gen_view(g, &lp, nullptr, nullptr, nullptr);
        expect(lp.nnz == 3, "nonzeros clamped to the rows");
        features(lp, 0, 0, "one-column LP");
        pm_gen_free(g);
    }
    // rolling-horizon windows
    for (double noise : {0.0, 0.2}) {
        pm_rh_params p;
        pm_rh_params_default(&p);
        p.demand_noise = noise;
        p.blend_noise = noise / 2;
        p.perturb_seed = 11;
        pm_rh* h = nullptr;
        expect(pm_rh_create(&p, &h) == 0, "plan created");
        int32_t rpp = 0, cpp = 0, windows = 0;
        expect(pm_rh_window_size(h, &rpp, &cpp, &windows) == 0 && rpp == 215 && cpp == 400 && windows == 500,
               "window layout");
        for (int w : {0, 250, windows - 1}) {
            pm_gen_lp* g = nullptr;
            expect(pm_rh_window(h, w, &g) == 0, "window " + std::to_string(w));
            pm_lp lp;
            pm_gen_view(g, &lp, nullptr, nullptr, nullptr);
            expect(lp.m == 24 * rpp && lp.n == 24 * cpp, "window size");
            features(lp, cpp, rpp, "window " + std::to_string(w));
            pm_gen_free(g);
        }
        pm_gen_lp* g = nullptr;
        expect(pm_rh_window(h, windows, &g) != 0 && g == nullptr, "window past the end rejected");
        pm_rh_free(h);
    }
    std::printf("%s\n", failures == 0 ? "ALL PASSED" : "FAILED");
    return failures == 0 ? 0 : 1;
}
