// ref_driver.cpp -- command-line driver around the reference's own CPU
// solver, compiled from the reference sources by oracle/Makefile into
// oracle/_ref/ref_driver.  TEST INFRASTRUCTURE: it gives tests FP64
// objectives and infeasible/unbounded verdicts of the reference's
// RevisedSimplex (src/simplex/revised_simplex.cpp:484-680) on small LPs.
//
//   ref_driver simplex <model.mps> [x.out] -> "status objective iterations"
//                                    (and x, one value per line, to x.out)
//   ref_driver ruiz <model.mps> <passes> [out.txt] -> "sum_row_scale sum_col_scale sum|A|"
//       (and to out.txt, %.17g one per line: row_scale, col_scale, then per
//       row its (column, value) entries, c, row bounds, variable bounds)
//   ref_driver cert <model.mps> <x.txt> <status> <tolerance> <out.cert>
//       the reference's generate_certificate + write_certificate
//       (src/certificate/certificate.cpp:23-153) for the given x.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <string>
#include <vector>

#include "pramaan/certificate.hpp"

#include "pramaan/mps_parser.hpp"
#include "pramaan/presolve.hpp"
#include "pramaan/simplex.hpp"
#include "pramaan/transformation_ledger.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: ref_driver simplex|ruiz <model.mps> [passes]\n");
        return 2;
    }
    try {
        pramaan::ModelIR model = pramaan::parse_mps(argv[2]);
        if (std::strcmp(argv[1], "simplex") == 0) {
            pramaan::RevisedSimplex solver;
            const pramaan::SolveResult r = solver.solve(model);
            std::printf("%d %.17g %d\n", static_cast<int>(r.status), r.objective_value, r.iterations);
            if (argc > 3) {
                std::ofstream out(argv[3]);
                out << std::setprecision(std::numeric_limits<double>::max_digits10);
                for (double v : r.x) out << v << "\n";
            }
            return 0;
        }
        if (std::strcmp(argv[1], "ruiz") == 0) {
            const int passes = argc > 3 ? std::atoi(argv[3]) : 10;
            pramaan::TransformationLedger ledger;
            const pramaan::ScalingFactors f = pramaan::ruiz_scale(model, ledger, passes);
            double sr = 0, sc = 0, sa = 0;
            for (double v : f.row_scale) sr += v;
            for (double v : f.col_scale) sc += v;
            for (double v : model.A.values()) sa += std::fabs(v);
            std::printf("%.17g %.17g %.17g\n", sr, sc, sa);
            if (argc > 4) {
                std::FILE* out = std::fopen(argv[4], "w");
                if (!out) return 1;
                for (double v : f.row_scale) std::fprintf(out, "%.17g\n", v);
                for (double v : f.col_scale) std::fprintf(out, "%.17g\n", v);
                const auto& rp = model.A.rowPtr();
                for (int r = 0; r < model.numRows(); ++r)
                    for (int k = rp[r]; k < rp[r + 1]; ++k)
                        std::fprintf(out, "%d %.17g\n", model.A.colIdx()[k], model.A.values()[k]);
                for (double v : model.obj_coeffs) std::fprintf(out, "%.17g\n", v);
                for (int r = 0; r < model.numRows(); ++r)
                    std::fprintf(out, "%.17g %.17g\n", model.row_lower[r], model.row_upper[r]);
                for (int j = 0; j < model.numVars(); ++j)
                    std::fprintf(out, "%.17g %.17g\n", model.var_lower[j], model.var_upper[j]);
                std::fclose(out);
            }
            return 0;
        }
        if (std::strcmp(argv[1], "cert") == 0 && argc > 6) {
            std::vector<double> x;
            std::ifstream in(argv[3]);
            for (double v; in >> v;) x.push_back(v);
            pramaan::Certificate c = pramaan::generate_certificate(model, x, argv[4], 0);
            c.tolerance = std::strtod(argv[5], nullptr);
            pramaan::write_certificate(c, argv[6]);
            return 0;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 2;
}
