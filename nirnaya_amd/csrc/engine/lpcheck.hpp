// Validation of a caller's LP (every index range-checked once on the host:
// the kernels trust them) and the counting transpose CSC(A) -> CSR(A).
#pragma once
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "../../../include/nirnaya/pdhg.h"

namespace nir {
namespace gpu {

inline void check_lp(const nir_lp* lp) {
    if (!lp) throw std::invalid_argument("null LP");
    if (lp->nrow < 0 || lp->ncol < 0) throw std::invalid_argument("negative dimensions");
    if (lp->ncol > 0 && (!lp->col_start || !lp->cost || !lp->collo || !lp->colup))
        throw std::invalid_argument("missing column arrays");
    if (lp->nrow > 0 && (!lp->rowlo || !lp->rowup)) throw std::invalid_argument("missing row arrays");
    const int64_t nnz = lp->ncol > 0 ? lp->col_start[lp->ncol] : 0;
    if (lp->ncol > 0 && lp->col_start[0] != 0) throw std::invalid_argument("col_start[0] != 0");
    if (nnz < 0 || nnz >= (int64_t(1) << 31) - 1) throw std::invalid_argument("nnz out of int32 range");
    if (nnz > 0 && (!lp->row_index || !lp->value)) throw std::invalid_argument("missing matrix arrays");
    for (int32_t j = 0; j < lp->ncol; ++j)
        if (lp->col_start[j + 1] < lp->col_start[j]) throw std::invalid_argument("col_start not non-decreasing");
    for (int64_t p = 0; p < nnz; ++p)
        if (lp->row_index[p] < 0 || lp->row_index[p] >= lp->nrow) throw std::invalid_argument("row index out of range");
}

// Most nonzeros in one row and in one column of A (row counts by a pass over
// the row indices).
inline int64_t longest_line(const nir_lp* lp) {
    int64_t most = 0;
    for (int32_t j = 0; j < lp->ncol; ++j) most = std::max<int64_t>(most, lp->col_start[j + 1] - lp->col_start[j]);
    std::vector<int64_t> rc(static_cast<size_t>(std::max(lp->nrow, 0)), 0);
    const int64_t nnz = lp->ncol > 0 ? lp->col_start[lp->ncol] : 0;
    for (int64_t p = 0; p < nnz; ++p) most = std::max(most, ++rc[lp->row_index[p]]);
    return most;
}

// CSR of A from the caller's CSC by a counting transpose; column indices come
// out ascending within each row (the order of the sequential row sums).
inline void csr_from_csc(const nir_lp* lp, std::vector<int>& ptr, std::vector<int>& col, std::vector<double>& val) {
    const int m = lp->nrow, n = lp->ncol;
    const int nnz = n > 0 ? static_cast<int>(lp->col_start[n]) : 0;
    ptr.assign(static_cast<size_t>(m) + 1, 0);
    col.assign(static_cast<size_t>(nnz), 0);
    val.assign(static_cast<size_t>(nnz), 0.0);
    for (int q = 0; q < nnz; ++q) ++ptr[lp->row_index[q] + 1];
    for (int i = 0; i < m; ++i) ptr[i + 1] += ptr[i];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int j = 0; j < n; ++j)
        for (int64_t q = lp->col_start[j]; q < lp->col_start[j + 1]; ++q) {
            const int r = fill[lp->row_index[q]]++;
            col[r] = j;
            val[r] = lp->value[q];
        }
}

}  // namespace gpu
}  // namespace nir
