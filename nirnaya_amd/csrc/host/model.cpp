#include "model.hpp"

#include <algorithm>
#include <utility>

namespace nir {

CscMatrix CscMatrix::assemble(Idx nrow, Idx ncol, const std::vector<Entry>& entries) {
    auto keep = [&](const Entry& e) {
        return e.row >= 0 && e.row < nrow && e.col >= 0 && e.col < ncol && e.val != 0.0;
    };
    // Bucket the entries by column, preserving their input order.
    std::vector<Off> bucket(static_cast<size_t>(ncol) + 1, 0);
    for (const Entry& e : entries)
        if (keep(e)) ++bucket[static_cast<size_t>(e.col) + 1];
    for (Idx j = 0; j < ncol; ++j) bucket[j + 1] += bucket[j];
    std::vector<std::pair<Idx, double>> cells(static_cast<size_t>(bucket[ncol]));
    {
        std::vector<Off> fill(bucket.begin(), bucket.end() - 1);
        for (const Entry& e : entries)
            if (keep(e)) cells[fill[e.col]++] = {e.row, e.val};
    }
    CscMatrix M(nrow, ncol);
    M.index.reserve(cells.size());
    M.value.reserve(cells.size());
    for (Idx j = 0; j < ncol; ++j) {
        M.start[j] = static_cast<Off>(M.index.size());
        auto first = cells.begin() + bucket[j];
        auto last = cells.begin() + bucket[j + 1];
        if (first == last) continue;
        // Row-only ordering: the order among duplicates of one row is the
        // sort's, and their sum is taken in that order.
        std::sort(first, last, [](const std::pair<Idx, double>& a, const std::pair<Idx, double>& b) {
            return a.first < b.first;
        });
        Idx row = first->first;
        double acc = first->second;
        for (auto it = first + 1; it != last; ++it) {
            if (it->first == row) {
                acc += it->second;
                continue;
            }
            if (acc != 0.0) {
                M.index.push_back(row);
                M.value.push_back(acc);
            }
            row = it->first;
            acc = it->second;
        }
        if (acc != 0.0) {
            M.index.push_back(row);
            M.value.push_back(acc);
        }
    }
    M.start[ncol] = static_cast<Off>(M.index.size());
    return M;
}

void CscMatrix::multiply(const std::vector<double>& x, std::vector<double>& y) const {
    y.assign(static_cast<size_t>(nrow), 0.0);
    for (Idx j = 0; j < ncol; ++j) {
        const double xj = x[j];
        if (xj == 0.0) continue;
        for (Off p = start[j]; p < start[j + 1]; ++p) y[index[p]] += value[p] * xj;
    }
}

void CscMatrix::multiply_transpose(const std::vector<double>& x, std::vector<double>& y) const {
    y.assign(static_cast<size_t>(ncol), 0.0);
    for (Idx j = 0; j < ncol; ++j) {
        double acc = 0.0;
        for (Off p = start[j]; p < start[j + 1]; ++p) acc += value[p] * x[index[p]];
        y[j] = acc;
    }
}

std::string CscMatrix::check() const {
    if (static_cast<Idx>(start.size()) != ncol + 1) return "start.size() != ncol + 1";
    if (start[0] != 0) return "start[0] != 0";
    for (Idx j = 0; j < ncol; ++j) {
        if (start[j] > start[j + 1]) return "start not non-decreasing";
        Idx prev = -1;
        for (Off p = start[j]; p < start[j + 1]; ++p) {
            if (index[p] < 0 || index[p] >= nrow) return "row index out of bounds";
            if (index[p] <= prev) return "row indices not strictly ascending";
            if (value[p] == 0.0) return "explicit zero stored";
            prev = index[p];
        }
    }
    return "";
}

CsrMatrix CsrMatrix::transpose_of(const CscMatrix& A) {
    CsrMatrix R;
    R.nrow = A.nrow;
    R.ncol = A.ncol;
    R.start.assign(static_cast<size_t>(A.nrow) + 1, 0);
    const Off nz = A.nnz();
    for (Off p = 0; p < nz; ++p) ++R.start[static_cast<size_t>(A.index[p]) + 1];
    for (Idx i = 0; i < A.nrow; ++i) R.start[i + 1] += R.start[i];
    R.index.resize(static_cast<size_t>(nz));
    R.value.resize(static_cast<size_t>(nz));
    std::vector<Off> fill(R.start.begin(), R.start.end() - 1);
    for (Idx j = 0; j < A.ncol; ++j) {
        for (Off p = A.start[j]; p < A.start[j + 1]; ++p) {
            Off q = fill[A.index[p]]++;
            R.index[q] = j;
            R.value[q] = A.value[p];
        }
    }
    return R;
}

CscMatrix CsrMatrix::to_csc() const {
    CscMatrix C(nrow, ncol);
    const Off nz = start.empty() ? 0 : start[nrow];
    for (Off p = 0; p < nz; ++p) ++C.start[static_cast<size_t>(index[p]) + 1];
    for (Idx j = 0; j < ncol; ++j) C.start[j + 1] += C.start[j];
    C.index.resize(static_cast<size_t>(nz));
    C.value.resize(static_cast<size_t>(nz));
    std::vector<Off> fill(C.start.begin(), C.start.end() - 1);
    for (Idx i = 0; i < nrow; ++i) {
        for (Off p = start[i]; p < start[i + 1]; ++p) {
            Off q = fill[index[p]]++;
            C.index[q] = i;
            C.value[q] = value[p];
        }
    }
    return C;
}

bool Model::is_mip() const {
    for (VarType t : vartype)
        if (t == VarType::Integer || t == VarType::Binary) return true;
    return false;
}

Idx Model::num_integers() const {
    Idx n = 0;
    for (VarType t : vartype)
        if (t == VarType::Integer || t == VarType::Binary) ++n;
    return n;
}

double Model::eval_objective(const std::vector<double>& x) const {
    DD acc(obj_offset);
    for (Idx j = 0; j < ncol; ++j)
        if (cost[j] != 0.0 && x[j] != 0.0) acc.add_prod(cost[j], x[j]);
    if (Q.nnz() > 0) {
        DD q(0.0);
        for (Idx j = 0; j < Q.ncol; ++j) {
            const double xj = x[j];
            if (xj == 0.0) continue;
            for (Off p = Q.start[j]; p < Q.start[j + 1]; ++p) {
                const Idx i = Q.index[p];
                if (i == j) q.add_prod(0.5 * Q.value[p], xj * xj);
                else if (i < j) q.add_prod(Q.value[p], x[i] * xj);
            }
        }
        acc.add(q);
    }
    return acc.value();
}

Model Model::to_internal_min() const {
    Model m = *this;
    if (sense == Sense::Maximize) {
        m.sense = Sense::Minimize;
        m.obj_offset = -obj_offset;
        for (double& c : m.cost) c = -c;
        for (double& v : m.Q.value) v = -v;
    }
    return m;
}

Status Model::validate(std::string& message) {
    const size_t n = static_cast<size_t>(ncol), m = static_cast<size_t>(nrow);
    if (cost.size() != n || collo.size() != n || colup.size() != n || vartype.size() != n) {
        message = "Column array dimension mismatch";
        return Status::InputError;
    }
    if (rowlo.size() != m || rowup.size() != m) {
        message = "Row array dimension mismatch";
        return Status::InputError;
    }
    for (Idx j = 0; j < ncol; ++j) {
        if (!std::isfinite(cost[j])) {
            message = "Non-finite objective coefficient on column " + std::to_string(j);
            return Status::InputError;
        }
        collo[j] = clean_lower(collo[j]);
        colup[j] = clean_upper(colup[j]);
        const bool integral = vartype[j] == VarType::Integer || vartype[j] == VarType::Binary;
        if (vartype[j] == VarType::Binary) {
            collo[j] = amax(0.0, collo[j]);
            colup[j] = amin(1.0, colup[j]);
        }
        if (integral) {
            if (finite_bound(collo[j])) collo[j] = std::ceil(collo[j] - 1e-9);
            if (finite_bound(colup[j])) colup[j] = std::floor(colup[j] + 1e-9);
        }
        if (collo[j] > colup[j]) {
            // Crossings within 1e-9 relative are snapped, larger ones are infeasible.
            if (collo[j] <= colup[j] + 1e-9 * (1.0 + std::fabs(colup[j]))) {
                collo[j] = colup[j];
            } else {
                message = "Infeasible column bounds on variable " +
                          (j < static_cast<Idx>(col_names.size()) ? col_names[j] : std::to_string(j));
                return Status::Infeasible;
            }
        }
        if (vartype[j] == VarType::Integer && collo[j] >= 0.0 && colup[j] <= 1.0)
            vartype[j] = VarType::Binary;
    }
    for (Idx i = 0; i < nrow; ++i) {
        rowlo[i] = clean_lower(rowlo[i]);
        rowup[i] = clean_upper(rowup[i]);
        if (rowlo[i] > rowup[i]) {
            if (rowlo[i] <= rowup[i] + 1e-9 * (1.0 + std::fabs(rowup[i]))) {
                rowlo[i] = rowup[i];
            } else {
                message = "Infeasible row bounds on row " +
                          (i < static_cast<Idx>(row_names.size()) ? row_names[i] : std::to_string(i));
                return Status::Infeasible;
            }
        }
    }
    for (double v : A.value) {
        if (!std::isfinite(v)) {
            message = "Non-finite matrix entry";
            return Status::InputError;
        }
    }
    return Status::Optimal;
}

Idx ModelBuilder::add_col(double cost, double lo, double up, VarType t, const std::string& name) {
    const Idx j = m_.ncol++;
    m_.cost.push_back(cost);
    m_.collo.push_back(lo);
    m_.colup.push_back(up);
    m_.vartype.push_back(t);
    m_.col_names.push_back(name.empty() ? "C" + std::to_string(j) : name);
    return j;
}

Idx ModelBuilder::add_row(double lo, double up, const Idx* idx, const double* val, int nnz,
                          const std::string& name) {
    const Idx i = m_.nrow++;
    m_.rowlo.push_back(lo);
    m_.rowup.push_back(up);
    m_.row_names.push_back(name.empty() ? "R" + std::to_string(i) : name);
    for (int k = 0; k < nnz; ++k)
        if (val[k] != 0.0) entries_.push_back({i, idx[k], val[k]});
    return i;
}

void ModelBuilder::add_quad(Idx i, Idx j, double v) {
    if (v == 0.0) return;
    if (i > j) std::swap(i, j);
    qentries_.push_back({i, j, v});
}

Model ModelBuilder::finalize() const {
    Model m = m_;
    m.A = CscMatrix::assemble(m.nrow, m.ncol, entries_);
    m.Q = qentries_.empty() ? CscMatrix(m.ncol, m.ncol) : CscMatrix::assemble(m.ncol, m.ncol, qentries_);
    return m;
}

}  // namespace nir
