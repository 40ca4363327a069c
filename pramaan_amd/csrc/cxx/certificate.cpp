// certificate.cpp -- certificates V1 (the reference's format) and V2 (with
// the FP64 dual), writer, reader and checker; see certificate.hpp.
#include "pramaan/certificate.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace pramaan {
namespace {

struct Primal {
    double objective, row_viol, bound_viol, int_viol;
};

// Objective (model sense, with offset) and the three primal violations,
// in the reference's order of operations (certificate.cpp:37-89).
Primal primal_measures(const ModelIR& m, const std::vector<double>& x) {
    Primal p{m.obj_offset, 0.0, 0.0, 0.0};
    const int n = m.numVars(), rows = m.numRows();
    for (int j = 0; j < n; ++j) p.objective += m.obj_coeffs[j] * x[j];
    const std::vector<double> act = m.A.multiply(x);
    for (int r = 0; r < rows; ++r) {
        if (m.row_lower[r] > -kInfinity && m.row_lower[r] - act[r] > p.row_viol) p.row_viol = m.row_lower[r] - act[r];
        if (m.row_upper[r] < kInfinity && act[r] - m.row_upper[r] > p.row_viol) p.row_viol = act[r] - m.row_upper[r];
    }
    for (int j = 0; j < n; ++j) {
        if (m.var_lower[j] > -kInfinity && m.var_lower[j] - x[j] > p.bound_viol) p.bound_viol = m.var_lower[j] - x[j];
        if (m.var_upper[j] < kInfinity && x[j] - m.var_upper[j] > p.bound_viol) p.bound_viol = x[j] - m.var_upper[j];
    }
    for (int j = 0; j < n; ++j)
        if (m.var_types[j] == VarType::kInteger) {
            const double v = std::abs(x[j] - std::round(x[j]));
            if (v > p.int_viol) p.int_viol = v;
        }
    return p;
}

struct Dual {
    double dual_residual, complementarity, gap;
};

// FP64 dual measures of (x, y), y in the reference sign (minimisation form
// cost c_min = +-c, reduced cost rc = c_min + A^T y, row multiplier t = -y):
//  dual residual   max over columns of the part of rc whose sign the
//                  finite bounds of x_j do not allow, and over rows of the
//                  part of t whose sign the finite row bounds do not allow;
//  complementarity max over |lambda_j (x_j - bound)| and |t_i (A x - bound)|
//                  with the bound the multiplier's sign selects;
//  gap             |c_min^T x - (sum t_i bound_i + sum lambda_j bound_j)|.
Dual dual_measures(const ModelIR& m, const std::vector<double>& x, const std::vector<double>& y) {
    const int n = m.numVars(), rows = m.numRows();
    const double sg = m.obj_sense == ObjSense::kMaximize ? -1.0 : 1.0;
    std::vector<double> rc(n, 0.0);
    const auto& rp = m.A.rowPtr();
    const auto& ci = m.A.colIdx();
    const auto& v = m.A.values();
    for (int r = 0; r < rows; ++r)
        for (int k = rp[r]; k < rp[r + 1]; ++k) rc[ci[k]] += v[k] * y[r];
    const std::vector<double> act = m.A.multiply(x);
    Dual d{0.0, 0.0, 0.0};
    double pobj = 0.0, dobj = 0.0;
    for (int j = 0; j < n; ++j) {
        const double c = sg * m.obj_coeffs[j];
        pobj += c * x[j];
        const double red = c + rc[j];
        const bool lf = m.var_lower[j] > -kInfinity, uf = m.var_upper[j] < kInfinity;
        double lam = red;
        if (!lf && lam > 0.0) lam = 0.0;
        if (!uf && lam < 0.0) lam = 0.0;
        d.dual_residual = std::max(d.dual_residual, std::abs(red - lam));
        if (lam > 0.0) {
            dobj += lam * m.var_lower[j];
            d.complementarity = std::max(d.complementarity, std::abs(lam * (x[j] - m.var_lower[j])));
        } else if (lam < 0.0) {
            dobj += lam * m.var_upper[j];
            d.complementarity = std::max(d.complementarity, std::abs(lam * (x[j] - m.var_upper[j])));
        }
    }
    for (int r = 0; r < rows; ++r) {
        const double t = -y[r];
        const bool lf = m.row_lower[r] > -kInfinity, uf = m.row_upper[r] < kInfinity;
        double tp = t;
        if (!lf && tp > 0.0) tp = 0.0;
        if (!uf && tp < 0.0) tp = 0.0;
        d.dual_residual = std::max(d.dual_residual, std::abs(t - tp));
        if (tp > 0.0) {
            dobj += tp * m.row_lower[r];
            d.complementarity = std::max(d.complementarity, std::abs(tp * (act[r] - m.row_lower[r])));
        } else if (tp < 0.0) {
            dobj += tp * m.row_upper[r];
            d.complementarity = std::max(d.complementarity, std::abs(tp * (act[r] - m.row_upper[r])));
        }
    }
    d.gap = std::abs(pobj - dobj);
    return d;
}

std::string fmt(double v) {
    std::ostringstream s;
    s << std::setprecision(std::numeric_limits<double>::max_digits10) << v;
    return s.str();
}

}  // namespace

Certificate generate_certificate(const ModelIR& model, const std::vector<double>& x, const std::string& status,
                                 int ledger_entries) {
    if (static_cast<int>(x.size()) != model.numVars())
        throw std::invalid_argument("certificate: x has " + std::to_string(x.size()) + " entries for " +
                                    std::to_string(model.numVars()) + " columns");
    Certificate c;
    c.status = status;
    c.num_rows = model.numRows();
    c.num_cols = model.numVars();
    c.model_fingerprint = std::to_string(computeModelFingerprint(model));
    c.ledger_entries = ledger_entries;
    c.x = x;
    const Primal p = primal_measures(model, x);
    c.objective_value = p.objective;
    c.primal_residual = p.row_viol;
    c.bound_violation = p.bound_viol;
    c.integrality_violation = p.int_viol;
    return c;
}

Certificate generate_certificate_kkt(const ModelIR& model, const std::vector<double>& x,
                                     const std::vector<double>& y, const std::string& status, int ledger_entries,
                                     double tolerance) {
    if (static_cast<int>(y.size()) != model.numRows())
        throw std::invalid_argument("certificate: y has " + std::to_string(y.size()) + " entries for " +
                                    std::to_string(model.numRows()) + " rows");
    Certificate c = generate_certificate(model, x, status, ledger_entries);
    c.format_id = "PRAMAAN_CERTIFICATE_V2";
    c.tolerance = tolerance;
    c.y = y;
    const Dual d = dual_measures(model, x, y);
    c.dual_residual = d.dual_residual;
    c.complementarity_residual = d.complementarity;
    c.objective_bound_gap = d.gap;
    return c;
}

// Field order, key spelling and max_digits10 output follow the reference
// writer (src/certificate/certificate.cpp:113-153); V2 adds Y ... END_Y.
void write_certificate(const Certificate& c, const std::string& path) {
    std::ofstream out(path);
    if (!out.is_open()) throw std::runtime_error("certificate: " + path + " could not be created");
    out << std::setprecision(std::numeric_limits<double>::max_digits10);
    auto opt = [&](const char* key, const std::optional<double>& v) {
        out << key << " ";
        if (v) out << *v << "\n";
        else out << "UNAVAILABLE\n";
    };
    out << c.format_id << "\n"
        << "STATUS " << c.status << "\n"
        << "ROWS " << c.num_rows << "\n"
        << "COLS " << c.num_cols << "\n"
        << "OBJECTIVE " << c.objective_value << "\n"
        << "TOLERANCE " << c.tolerance << "\n"
        << "PRIMAL_RESIDUAL " << c.primal_residual << "\n";
    opt("DUAL_RESIDUAL", c.dual_residual);
    opt("COMPLEMENTARITY_RESIDUAL", c.complementarity_residual);
    out << "BOUND_VIOLATION " << c.bound_violation << "\n"
        << "INTEGRALITY_VIOLATION " << c.integrality_violation << "\n";
    opt("OBJECTIVE_BOUND_GAP", c.objective_bound_gap);
    out << "MODEL_FINGERPRINT " << c.model_fingerprint << "\n"
        << "LEDGER_ENTRIES " << c.ledger_entries << "\n"
        << "X\n";
    for (double v : c.x) out << v << "\n";
    out << "END_X\n";
    if (c.format_id == "PRAMAAN_CERTIFICATE_V2") {
        out << "Y\n";
        for (double v : c.y) out << v << "\n";
        out << "END_Y\n";
    }
    out << "END_CERTIFICATE\n";
    if (!out.good()) throw std::runtime_error("certificate: writing " + path + " failed");
}

// The reference reader's grammar and rejections (certificate.cpp:159-287):
// CRLF tolerated, NaN/Inf/trailing garbage rejected, x (and y) sizes must
// match COLS (ROWS), nothing but blank lines after END_CERTIFICATE.
Certificate read_certificate(const std::string& path) {
    std::ifstream in(path);
    const auto bad = [&path](const std::string& why) { throw std::runtime_error("certificate " + path + ": " + why); };
    if (!in.is_open()) bad("cannot be read");
    std::string line;
    auto next = [&]() -> std::string {
        if (!std::getline(in, line)) bad("ends before END_CERTIFICATE");
        if (!line.empty() && line.back() == '\r') line.pop_back();
        return line;
    };
    auto value = [&](const std::string& l, const std::string& key) {
        if (l.compare(0, key.size(), key) != 0) bad("line \"" + l + "\" where the " + key + "field belongs");
        return l.substr(key.size());
    };
    auto num = [&](const std::string& s, const std::string& key) {
        if (s.empty()) bad(key + "has no value");
        size_t pos = 0;
        const double d = std::stod(s, &pos);
        while (pos < s.size() && std::isspace(static_cast<unsigned char>(s[pos]))) ++pos;
        if (pos != s.size()) bad(key + "value \"" + s + "\" is not a number alone");
        if (!std::isfinite(d)) bad(key + "value \"" + s + "\" is not finite");
        return d;
    };
    auto integer = [&](const std::string& s, const std::string& key) {
        if (s.empty()) bad(key + "has no value");
        size_t pos = 0;
        const int i = std::stoi(s, &pos);
        while (pos < s.size() && std::isspace(static_cast<unsigned char>(s[pos]))) ++pos;
        if (pos != s.size()) bad(key + "value \"" + s + "\" is not an integer alone");
        return i;
    };
    auto optional = [&](const std::string& s, const std::string& key) -> std::optional<double> {
        if (s == "UNAVAILABLE" || s == "NOT_COMPUTED") return std::nullopt;
        return num(s, key);
    };
    Certificate c;
    if (!std::getline(in, line)) bad("is empty");
    if (!line.empty() && line.back() == '\r') line.pop_back();
    c.format_id = line;
    if (c.format_id != "PRAMAAN_CERTIFICATE_V1" && c.format_id != "PRAMAAN_CERTIFICATE_V2")
        bad("format \"" + c.format_id + "\" is neither V1 nor V2");
    c.status = value(next(), "STATUS ");
    c.num_rows = integer(value(next(), "ROWS "), "ROWS ");
    c.num_cols = integer(value(next(), "COLS "), "COLS ");
    if (c.num_rows < 0 || c.num_cols < 0) bad("ROWS and COLS must not be negative");
    c.objective_value = num(value(next(), "OBJECTIVE "), "OBJECTIVE ");
    c.tolerance = num(value(next(), "TOLERANCE "), "TOLERANCE ");
    if (c.tolerance <= 0.0) bad("TOLERANCE must be positive");
    c.primal_residual = num(value(next(), "PRIMAL_RESIDUAL "), "PRIMAL_RESIDUAL ");
    c.dual_residual = optional(value(next(), "DUAL_RESIDUAL "), "DUAL_RESIDUAL ");
    c.complementarity_residual = optional(value(next(), "COMPLEMENTARITY_RESIDUAL "), "COMPLEMENTARITY_RESIDUAL ");
    c.bound_violation = num(value(next(), "BOUND_VIOLATION "), "BOUND_VIOLATION ");
    c.integrality_violation = num(value(next(), "INTEGRALITY_VIOLATION "), "INTEGRALITY_VIOLATION ");
    c.objective_bound_gap = optional(value(next(), "OBJECTIVE_BOUND_GAP "), "OBJECTIVE_BOUND_GAP ");
    c.model_fingerprint = value(next(), "MODEL_FINGERPRINT ");
    c.ledger_entries = integer(value(next(), "LEDGER_ENTRIES "), "LEDGER_ENTRIES ");
    if (c.ledger_entries < 0) bad("LEDGER_ENTRIES must not be negative");
    if (next() != "X") bad("line \"" + line + "\" where the X block starts");
    for (;;) {
        const std::string l = next();
        if (l == "END_X") break;
        c.x.push_back(num(l, "X element"));
    }
    if (static_cast<int>(c.x.size()) != c.num_cols) bad("X has " + std::to_string(c.x.size()) + " values, COLS says " + std::to_string(c.num_cols));
    if (c.format_id == "PRAMAAN_CERTIFICATE_V2") {
        if (next() != "Y") bad("line \"" + line + "\" where the Y block starts");
        for (;;) {
            const std::string l = next();
            if (l == "END_Y") break;
            c.y.push_back(num(l, "Y element"));
        }
        if (static_cast<int>(c.y.size()) != c.num_rows) bad("Y has " + std::to_string(c.y.size()) + " values, ROWS says " + std::to_string(c.num_rows));
    }
    if (next() != "END_CERTIFICATE") bad("line \"" + line + "\" where END_CERTIFICATE belongs");
    std::string extra;
    if (std::getline(in, extra) && extra.find_first_not_of(" \r\n\t") != std::string::npos)
        bad("text follows END_CERTIFICATE");
    return c;
}

// Checks 0-10 of the reference verifier (src/certificate/verifier_main.cpp:
// 68-221) in its order and with its tolerances; a V2 certificate replaces
// check 9 (dual fields must be UNAVAILABLE) with the recomputed dual checks.
VerifyResult verify_certificate(const ModelIR& model, const Certificate& c) {
    VerifyResult res;
    auto fail = [&](const std::string& why) { res.failures.push_back(why); };
    const double tol = c.tolerance;
    const int m = model.numRows(), n = model.numVars();
    if (tol > 1e-4) fail("Tolerance too large: " + fmt(tol) + " > 1e-4");
    const bool v2 = c.format_id == "PRAMAAN_CERTIFICATE_V2";
    if (c.format_id != "PRAMAAN_CERTIFICATE_V1" && !v2) fail("Unknown certificate format: " + c.format_id);
    if (c.status != "OPTIMAL") fail("Certificate status is not OPTIMAL: " + c.status);
    if (c.num_rows != m) fail("Row count mismatch");
    if (c.num_cols != n) fail("Column count mismatch");
    if (static_cast<int>(c.x.size()) != n) {
        fail("Solution vector size mismatch");
        return res;
    }
    for (double v : c.x)
        if (!std::isfinite(v)) fail("x contains NaN or Inf");
    const Primal p = primal_measures(model, c.x);
    if (std::abs(p.objective - c.objective_value) > tol) fail("Objective mismatch: recomputed=" + fmt(p.objective));
    if (p.row_viol > tol) fail("Primal infeasible: max row violation=" + fmt(p.row_viol));
    if (std::abs(p.row_viol - c.primal_residual) > tol) fail("Primal residual mismatch");
    if (p.bound_viol > tol) fail("Bound violation=" + fmt(p.bound_viol));
    if (std::abs(p.bound_viol - c.bound_violation) > tol) fail("Bound violation mismatch");
    if (p.int_viol > tol) fail("Integrality violation=" + fmt(p.int_viol));
    if (std::abs(p.int_viol - c.integrality_violation) > tol) fail("Integrality violation mismatch");
    if (!v2) {
        if (c.dual_residual) fail("Dual residual is unsupported in V1, must be UNAVAILABLE");
        if (c.complementarity_residual) fail("Complementarity residual is unsupported in V1, must be UNAVAILABLE");
        if (c.objective_bound_gap) fail("Objective bound gap is unsupported in V1, must be UNAVAILABLE");
    } else {
        if (static_cast<int>(c.y.size()) != m) {
            fail("Dual vector size mismatch");
        } else if (!c.dual_residual || !c.complementarity_residual || !c.objective_bound_gap) {
            fail("V2 certificate must carry dual residual, complementarity and gap");
        } else {
            for (double v : c.y)
                if (!std::isfinite(v)) fail("y contains NaN or Inf");
            const Dual d = dual_measures(model, c.x, c.y);
            // The dual residual is judged against max(1, |c|_inf), the
            // normalisation of the reference's dual violation
            // (src/gpu/pdhg_solver.cu:142-157); complementarity and the gap
            // against 1 + |objective|.
            double cmax = 1.0;
            for (double v : model.obj_coeffs) cmax = std::max(cmax, std::abs(v));
            const double dscale = cmax, scale = 1.0 + std::abs(p.objective);
            if (d.dual_residual > tol * dscale) fail("Dual infeasible: " + fmt(d.dual_residual));
            if (std::abs(d.dual_residual - *c.dual_residual) > tol * dscale) fail("Dual residual mismatch");
            if (d.complementarity > tol * scale) fail("Complementarity violated: " + fmt(d.complementarity));
            if (std::abs(d.complementarity - *c.complementarity_residual) > tol * scale)
                fail("Complementarity residual mismatch");
            if (d.gap > tol * scale) fail("Duality gap too large: " + fmt(d.gap));
            if (std::abs(d.gap - *c.objective_bound_gap) > tol * scale) fail("Objective bound gap mismatch");
        }
    }
    if (c.model_fingerprint != std::to_string(computeModelFingerprint(model))) fail("Model fingerprint mismatch");
    res.valid = res.failures.empty();
    return res;
}

}  // namespace pramaan

// ---- C entry points for tests and non-C++ hosts ----------------------------
#include "pramaan_cert.h"

namespace {
pramaan::ModelIR model_of(const pm_lp* lp) {
    using namespace pramaan;
    std::vector<CSRMatrix::Index> rp(lp->row_ptr, lp->row_ptr + lp->m + 1), ci(lp->col_idx, lp->col_idx + lp->nnz);
    std::vector<double> v(lp->values, lp->values + lp->nnz);
    return ModelIR(lp->sense == 1 ? ObjSense::kMaximize : ObjSense::kMinimize, lp->offset,
                   std::vector<double>(lp->c, lp->c + lp->n), CSRMatrix(rp, ci, v, lp->n),
                   std::vector<double>(lp->row_lower, lp->row_lower + lp->m),
                   std::vector<double>(lp->row_upper, lp->row_upper + lp->m), std::vector<std::string>(lp->m),
                   std::vector<double>(lp->var_lower, lp->var_lower + lp->n),
                   std::vector<double>(lp->var_upper, lp->var_upper + lp->n),
                   std::vector<VarType>(lp->n, VarType::kContinuous), std::vector<std::string>(lp->n));
}
}  // namespace

extern "C" {

// Writes a V1 certificate (y == NULL) or a V2 certificate for (x, y).
// Returns 0 on success, 1 on error (message in errbuf).
int pm_cert_write(const pm_lp* lp, const double* x, const double* y, const char* status, double tolerance,
                  const char* path, char* errbuf, int errlen) {
    try {
        const pramaan::ModelIR m = model_of(lp);
        std::vector<double> xv(x, x + lp->n);
        pramaan::Certificate c =
            y ? pramaan::generate_certificate_kkt(m, xv, std::vector<double>(y, y + lp->m), status, 0, tolerance)
              : pramaan::generate_certificate(m, xv, status, 0);
        c.tolerance = tolerance;
        pramaan::write_certificate(c, path);
        return 0;
    } catch (const std::exception& e) {
        if (errbuf && errlen > 0) std::snprintf(errbuf, static_cast<size_t>(errlen), "%s", e.what());
        return 1;
    }
}

// Verifies a certificate file against the LP: 0 valid, 1 invalid (reasons,
// one per line, in msg), 2 unreadable.
int pm_cert_verify(const pm_lp* lp, const char* path, char* msg, int msglen) {
    try {
        const pramaan::ModelIR m = model_of(lp);
        const pramaan::Certificate c = pramaan::read_certificate(path);
        const pramaan::VerifyResult r = pramaan::verify_certificate(m, c);
        std::string all;
        for (const auto& f : r.failures) all += f + "\n";
        if (msg && msglen > 0) std::snprintf(msg, static_cast<size_t>(msglen), "%s", all.c_str());
        return r.valid ? 0 : 1;
    } catch (const std::exception& e) {
        if (msg && msglen > 0) std::snprintf(msg, static_cast<size_t>(msglen), "%s", e.what());
        return 2;
    }
}

}  // extern "C"
