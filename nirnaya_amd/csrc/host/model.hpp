// LP model storage: compressed-column matrix, bounds, costs, names, and the
// incremental builder behind the C ABI.
//
// Semantics follow the reference:
//   src/core/csc_matrix.hpp:30-96   CSC assembly (per-column row sort,
//                                   duplicate summation, exact-zero drop)
//   src/core/csc_matrix.hpp:99-120  y = A x (column scatter) and y = A^T x
//   src/model/model.hpp:73-110      dd objective, max -> min conversion
//   src/model/model.hpp:117-178     validation and bound repair
//   src/model/model.hpp:181-241     ModelBuilder
#pragma once
#include <string>
#include <vector>

#include "core.hpp"

namespace nir {

struct Entry {
    Idx row;
    Idx col;
    double val;
};

struct CscMatrix {
    Idx nrow = 0;
    Idx ncol = 0;
    std::vector<Off> start{0};
    std::vector<Idx> index;
    std::vector<double> value;

    CscMatrix() = default;
    CscMatrix(Idx r, Idx c) : nrow(r), ncol(c), start(static_cast<size_t>(c) + 1, 0) {}

    Off nnz() const { return start.empty() ? 0 : start[ncol]; }

    // Canonical CSC from coordinate entries: out-of-range entries and exact
    // zeros are skipped, each column is sorted by row, duplicates summed in
    // sorted order, and sums that cancel to 0.0 are dropped.
    static CscMatrix assemble(Idx nrow, Idx ncol, const std::vector<Entry>& entries);

    // y = A x, accumulated column by column (columns with x_j == 0 skipped).
    void multiply(const std::vector<double>& x, std::vector<double>& y) const;
    // y = A^T x, one ascending-row dot product per column.
    void multiply_transpose(const std::vector<double>& x, std::vector<double>& y) const;
    // Empty string when canonical, else the first defect found.
    std::string check() const;
};

// Row-major copy (used by presolve and the device upload). Within a row the
// column indices come out ascending.
struct CsrMatrix {
    Idx nrow = 0;
    Idx ncol = 0;
    std::vector<Off> start{0};
    std::vector<Idx> index;
    std::vector<double> value;
    static CsrMatrix transpose_of(const CscMatrix& A);
    CscMatrix to_csc() const;
};

struct Model {
    std::string name = "unnamed";
    Sense sense = Sense::Minimize;
    double obj_offset = 0.0;
    Idx ncol = 0;
    Idx nrow = 0;
    std::vector<double> cost, collo, colup;
    std::vector<VarType> vartype;
    CscMatrix A;
    std::vector<double> rowlo, rowup;
    std::vector<std::string> col_names, row_names;
    CscMatrix Q;   // upper triangle of a quadratic objective (only its presence matters here)

    bool is_mip() const;
    bool is_qp() const { return Q.nnz() > 0; }
    Idx num_integers() const;

    // offset + c^T x (+ 0.5 x^T Q x) in double-double, skipping zero terms.
    double eval_objective(const std::vector<double>& x) const;
    Model to_internal_min() const;
    double user_objective_from_internal(double v) const {
        return sense == Sense::Maximize ? -v : v;
    }
    // Sanitises bounds in place; returns Optimal when valid, else the
    // failure status with `message` set.
    Status validate(std::string& message);
};

class ModelBuilder {
public:
    explicit ModelBuilder(const std::string& name = "model") { m_.name = name; }
    Idx add_col(double cost, double lo, double up, VarType t, const std::string& name);
    Idx add_row(double lo, double up, const Idx* idx, const double* val, int nnz,
                const std::string& name);
    void add_quad(Idx i, Idx j, double v);
    void set_sense(Sense s) { m_.sense = s; }
    void set_offset(double c0) { m_.obj_offset = c0; }
    Model finalize() const;

private:
    Model m_;
    std::vector<Entry> entries_;
    std::vector<Entry> qentries_;
};

}  // namespace nir
