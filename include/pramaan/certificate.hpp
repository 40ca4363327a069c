// certificate.hpp -- solution certificates and their independent checker.
//
// PRAMAAN_CERTIFICATE_V1 is the reference's text format
// (src/certificate/certificate.cpp:113-287): status, dimensions, FP64
// objective, tolerance, primal residual, dual / complementarity / gap
// (UNAVAILABLE in V1), bound and integrality violation, FNV-1a model
// fingerprint, ledger entries and x.  write_certificate produces the same
// bytes for the same fields, and verify_certificate applies the checks of
// the reference's pramaan-verify (src/certificate/verifier_main.cpp:58-229).
//
// PRAMAAN_CERTIFICATE_V2 adds the dual y (reference sign: reduced cost
// c + A^T y, minimisation form) after x and fills the three V1 gaps with
// FP64 values recomputed from (x, y): the dual residual (largest violation
// of the reduced-cost and row-multiplier signs the bounds allow), the
// complementarity residual (largest |multiplier x slack| over bounds and
// rows) and the objective bound gap |c^T x - dual objective|.  The checker
// recomputes them and accepts the dual residual up to tolerance*max(1,|c|inf)
// and complementarity and gap up to tolerance*(1+|objective|), and requires
// the written values to agree with the recomputed ones to the same bounds.
#pragma once

#include <optional>
#include <string>
#include <vector>

#include "pramaan/ir.hpp"

namespace pramaan {

struct Certificate {
    std::string format_id = "PRAMAAN_CERTIFICATE_V1";
    std::string status;
    int num_rows = 0;
    int num_cols = 0;
    double objective_value = 0.0;
    double tolerance = 1e-6;
    double primal_residual = 0.0;
    std::optional<double> dual_residual;
    std::optional<double> complementarity_residual;
    double bound_violation = 0.0;
    double integrality_violation = 0.0;
    std::optional<double> objective_bound_gap;
    std::string model_fingerprint;
    int ledger_entries = 0;
    std::vector<double> x;
    std::vector<double> y;  // V2 only
};

// V1: recomputes objective, primal residual, bound and integrality
// violation from the model and x (reference certificate.cpp:23-107).
Certificate generate_certificate(const ModelIR& model, const std::vector<double>& x, const std::string& status,
                                 int ledger_entries);
// V2: also takes the dual y and fills the dual fields in FP64.
Certificate generate_certificate_kkt(const ModelIR& model, const std::vector<double>& x,
                                     const std::vector<double>& y, const std::string& status, int ledger_entries,
                                     double tolerance);

void write_certificate(const Certificate& cert, const std::string& filepath);
Certificate read_certificate(const std::string& filepath);

struct VerifyResult {
    bool valid = false;
    std::vector<std::string> failures;
};
// The checks of the reference verifier, plus the dual checks for V2.
VerifyResult verify_certificate(const ModelIR& model, const Certificate& cert);

}  // namespace pramaan
