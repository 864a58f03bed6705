// Host-side steps on either side of the PDHG engine: geometric-mean
// equilibration rounded to powers of two, the presolve/postsolve pair, and the
// double-double solution checker.
//
//   scale_geometric  follows src/model/scaling.hpp:14-154
//   presolve         follows src/presolve/presolve.hpp:32-345 (rules R1-R7,
//                    R14, R20 in the same order and with the same tolerances)
//   verify_solution  follows src/io/sol_io.hpp:25-106
#pragma once
#include <string>
#include <vector>

#include "model.hpp"

namespace nir {

struct Scaling {
    bool active = false;
    std::vector<double> row;   // R (m): A' = R A C
    std::vector<double> col;   // C (n)

    // 2^round(log2 v), exponent clamped to [-30, 30]; 1 for v <= 0 or non-finite.
    static double to_pow2(double v);
    void unscale_primal(const std::vector<double>& xs, std::vector<double>& x) const;
    void unscale_duals(const std::vector<double>& ys, std::vector<double>& y) const;
};

// Returns the scaled copy of `orig` (or `orig` itself when scaling does not
// engage: disabled, no nonzeros, or max|a|/min|a| <= 10).
Model scale_geometric(const Model& orig, bool enable, Scaling& info);

struct Presolved {
    Status status = Status::Optimal;
    bool solved = false;          // nothing left to solve (or a verdict was reached)
    Model reduced;
    std::vector<Idx> col_map;     // reduced column -> original column
    std::vector<Idx> row_map;     // reduced row -> original row
    std::vector<double> fixed_x;  // values of removed columns (size: original ncol)
    Idx rows_removed = 0, cols_removed = 0, tightenings = 0;

    void postsolve(const std::vector<double>& x_red, std::vector<double>& x_orig) const;
};

Presolved presolve(const Model& orig, int level, int max_rounds);

struct CheckReport {
    bool passed = true;
    double max_bound_viol = 0.0;
    double max_row_viol = 0.0;
    double max_int_viol = 0.0;
    double computed_obj = 0.0;
    double obj_rel_err = 0.0;
    std::string worst_item;
};

// Bound, row (double-double activities) and integrality violations relative to
// 1 + |bound|, and the claimed objective against a double-double recompute.
CheckReport verify_solution(const Model& model, const std::vector<double>& x, double claimed_obj,
                            double tol);

// ".sol" text: "# status", "# objective", "# gap" header lines, then
// "<column name> <value>" per column (sol_io.hpp:110-170).
bool write_sol(const std::string& path, const Model& m, Status st, double obj, double gap,
               const std::vector<double>& x);
bool read_sol(const std::string& path, const Model& m, Status& st, double& obj,
              std::vector<double>& x);

// Free-format MPS reader/writer (mps_reader.hpp:15-413).
Status read_mps(const std::string& path, Model& out, std::string& err);
bool write_mps(const std::string& path, const Model& m);

// Binary LP container (".nlpb", lpb.cpp): the model's arrays as stored, with
// a checksum; the reader validates sizes, checksum and canonical form.
bool is_lpb_file(const std::string& path);
bool write_lpb(const std::string& path, const Model& m, std::string& err);
Status read_lpb(const std::string& path, Model& out, std::string& err);
// read_lpb for binary files (recognised by their magic), read_mps otherwise.
Status read_model(const std::string& path, Model& out, std::string& err);

}  // namespace nir
