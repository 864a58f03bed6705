// checker <model.mps> <solution.sol> [tol=1e-6]: independent double-double
// certification of a solution file (tools/checker.cpp:6-42 semantics; exit 0
// on PASS, 1 otherwise).
#include <cstdio>
#include <string>

#include "prep.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "Usage: checker <model.mps> <solution.sol> [tol=1e-6]\n");
        return 1;
    }
    double tol = 1e-6;
    if (argc >= 4) {
        try {
            tol = std::stod(argv[3]);
        } catch (const std::exception&) {
            std::fprintf(stderr, "FAIL: bad tolerance %s\n", argv[3]);
            return 1;
        }
    }
    nir::Model model;
    std::string err;
    if (nir::read_model(argv[1], model, err) != nir::Status::Optimal) {
        std::fprintf(stderr, "FAIL: Could not read model %s: %s\n", argv[1], err.c_str());
        return 1;
    }
    nir::Status claimed;
    double obj = 0.0;
    std::vector<double> x;
    if (!nir::read_sol(argv[2], model, claimed, obj, x)) {
        std::fprintf(stderr, "FAIL: Could not read solution %s\n", argv[2]);
        return 1;
    }
    const nir::CheckReport r = nir::verify_solution(model, x, obj, tol);
    std::printf("================================================================\n");
    std::printf("  Independent solution check (double-double)\n");
    std::printf("  Model:             %s (%d rows, %d cols)\n", model.name.c_str(), model.nrow, model.ncol);
    std::printf("  Claimed Objective: %.17g\n", obj);
    std::printf("  True Objective:    %.17g (rel err: %.3e)\n", r.computed_obj, r.obj_rel_err);
    std::printf("  Max Bound Viol:    %.3e\n", r.max_bound_viol);
    std::printf("  Max Row Viol:      %.3e\n", r.max_row_viol);
    std::printf("  Max Int Viol:      %.3e\n", r.max_int_viol);
    std::printf("  Verdict:           %s\n", r.passed ? "PASS" : ("FAIL (" + r.worst_item + ")").c_str());
    std::printf("================================================================\n");
    return r.passed ? 0 : 1;
}
