// Host-shell checks built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_host_sanitizers.py): the MPS and .nlpb readers and writers,
// validation, presolve/postsolve, geometric scaling and the double-double
// checker of nirnaya_amd/csrc/host, on a model file given on the command line
// and on seeded random models built through ModelBuilder (duplicate entries,
// zeros, infinite and fixed bounds, empty rows and columns). Any memory error
// or undefined behaviour aborts the run; the checks below make it exit 1 on a
// wrong result.
//
//   host_checks <model.mps> <tmpdir>
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../nirnaya_amd/csrc/host/prep.hpp"

namespace {

int failures = 0;
int exercised = 0;

void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what);
        ++failures;
    }
}

bool same_model(const nir::Model& a, const nir::Model& b) {
    return a.ncol == b.ncol && a.nrow == b.nrow && a.cost == b.cost && a.collo == b.collo &&
           a.colup == b.colup && a.rowlo == b.rowlo && a.rowup == b.rowup && a.A.start == b.A.start &&
           a.A.index == b.A.index && a.A.value == b.A.value;
}

// Every step of the host shell on one model.
void exercise(const nir::Model& model, const std::string& tmp, const std::string& tag) {
    nir::Model m = model;
    std::string msg;
    const nir::Status vs = m.validate(msg);
    if (vs != nir::Status::Optimal) return;   // rejected input: nothing more to run
    ++exercised;
    const nir::Model internal = m.to_internal_min();

    // .nlpb and MPS round trips.
    std::string err;
    const std::string lpb = tmp + "/" + tag + ".nlpb";
    check(nir::write_lpb(lpb, m, err), "write_lpb");
    nir::Model back;
    check(nir::read_lpb(lpb, back, err) == nir::Status::Optimal, "read_lpb");
    check(same_model(m, back), "nlpb round trip");
    const std::string mps = tmp + "/" + tag + ".mps";
    check(nir::write_mps(mps, m), "write_mps");
    nir::Model back2;
    check(nir::read_mps(mps, back2, err) == nir::Status::Optimal, "read_mps of written file");
    check(back2.ncol == m.ncol && back2.nrow == m.nrow && back2.A.nnz() == m.A.nnz(), "mps round trip shape");

    // Presolve and postsolve of a point inside the bounds.
    const nir::Presolved pre = nir::presolve(internal, 2, 20);
    if (!pre.solved) {
        std::vector<double> xr(static_cast<size_t>(pre.reduced.ncol), 0.0), xo;
        for (nir::Idx j = 0; j < pre.reduced.ncol; ++j) {
            const double l = pre.reduced.collo[j], u = pre.reduced.colup[j];
            xr[j] = std::isfinite(l) ? l : (std::isfinite(u) ? u : 0.0);
        }
        pre.postsolve(xr, xo);
        check(static_cast<nir::Idx>(xo.size()) == internal.ncol, "postsolve size");
    }

    // Geometric scaling: exact powers of two, and unscaling of scaled zeros.
    nir::Scaling sc;
    const nir::Model scaled = nir::scale_geometric(internal, true, sc);
    if (sc.active) {
        bool pow2 = true;
        for (double r : sc.row) pow2 = pow2 && std::ldexp(1.0, std::ilogb(r)) == r;
        for (double c : sc.col) pow2 = pow2 && std::ldexp(1.0, std::ilogb(c)) == c;
        check(pow2, "scaling factors are powers of two");
        std::vector<double> xs(static_cast<size_t>(scaled.ncol), 1.0), x;
        sc.unscale_primal(xs, x);
        check(x.size() == xs.size(), "unscale size");
    }

    // The checker on the bound midpoint with its own objective.
    std::vector<double> x(static_cast<size_t>(m.ncol), 0.0);
    for (nir::Idx j = 0; j < m.ncol; ++j) {
        const double l = m.collo[j], u = m.colup[j];
        x[j] = (std::isfinite(l) && std::isfinite(u)) ? 0.5 * (l + u) : (std::isfinite(l) ? l : (std::isfinite(u) ? u : 0.0));
    }
    const double obj = m.eval_objective(x);
    const nir::CheckReport rep = nir::verify_solution(m, x, obj, 1e-6);
    check(rep.obj_rel_err <= 1e-15, "checker objective of its own recompute");
}

nir::Model random_model(unsigned seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const int n = 5 + static_cast<int>(rng() % 60), m = 3 + static_cast<int>(rng() % 40);
    nir::ModelBuilder b("random");
    const double inf = INFINITY;
    for (int j = 0; j < n; ++j) {
        const int kind = static_cast<int>(rng() % 5);
        const double lo = kind == 0 ? -inf : (kind == 1 ? U(rng) : 0.0);
        const double up = kind == 2 ? inf : (kind == 3 ? lo : (lo > -inf ? lo + 1.0 + std::fabs(U(rng)) : 2.0));
        b.add_col(U(rng), lo, up, nir::VarType::Continuous, "c" + std::to_string(j));
    }
    for (int i = 0; i < m; ++i) {
        const int len = static_cast<int>(rng() % 6);   // 0 = an empty row
        std::vector<nir::Idx> idx;
        std::vector<double> val;
        for (int t = 0; t < len; ++t) {
            idx.push_back(static_cast<nir::Idx>(rng() % n));   // duplicates merge
            val.push_back(rng() % 7 == 0 ? 0.0 : U(rng));      // zeros are dropped
        }
        const int kind = static_cast<int>(rng() % 4);
        const double lo = kind == 0 ? -inf : U(rng);
        const double up = kind == 1 ? inf : (kind == 2 ? lo : (lo > -inf ? lo + std::fabs(U(rng)) : 1.0));
        b.add_row(lo, up, idx.data(), val.data(), len, "r" + std::to_string(i));
    }
    return b.finalize();
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: host_checks <model.mps> <tmpdir>\n");
        return 2;
    }
    nir::Model model;
    std::string err;
    if (nir::read_model(argv[1], model, err) != nir::Status::Optimal) {
        std::fprintf(stderr, "FAIL: could not read %s: %s\n", argv[1], err.c_str());
        return 1;
    }
    exercise(model, argv[2], "fixture");
    for (unsigned s = 1; s <= 200; ++s) exercise(random_model(s), argv[2], "random");
    std::printf("host checks: %d models through every step, %d failure(s)\n", exercised, failures);
    return failures == 0 ? 0 : 1;
}
