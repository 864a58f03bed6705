// Crossover hook: a simplex starting basis guessed from an (approximate)
// PDHG solution, in the layout the reference's simplex takes as a warm start
// (BasisSnapshot, src/lp/lp_solver.hpp:18-21, read by LpSolver::set_basis at
// :896-907): one status per variable - the ncol structural columns, then one
// logical per row whose value is the row activity (A x) and whose bounds are
// the row bounds (LpSolver::load, :75-104) - and the m basic variables.
//
// Rule: a variable is basic when it lies strictly inside its bounds (its
// distance to the nearer finite bound above tol * (1 + |bound|)); the rest sit
// at the bound they touch (Fixed when the bounds coincide, FreeZero for a
// free variable at 0). With more than m interior variables the m most
// interior (distance relative to 1 + |value|; free variables first) stay
// basic; with fewer, the logicals of the rows with the most slack are added
// until there are m. The reference's factorization repairs a singular guess
// by swapping logicals in (src/linalg/lu_factor.hpp:212-226).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/nirnaya/pdhg.h"
#include "core.hpp"

namespace nir {
void record_last_error(const std::string& what);
}

using nir::finite_bound;

extern "C" {

NIR_API int nir_basis_guess(const nir_lp* lp, const double* x, double tol, int8_t* status, int32_t* basic_index) {
    try {
        if (!lp || lp->nrow < 0 || lp->ncol < 0) throw std::invalid_argument("invalid LP");
        if (!status || (lp->nrow > 0 && !basic_index)) throw std::invalid_argument("missing output arrays");
        if (lp->ncol > 0 && (!x || !lp->col_start || !lp->collo || !lp->colup))
            throw std::invalid_argument("missing column arrays or point");
        if (lp->nrow > 0 && (!lp->rowlo || !lp->rowup)) throw std::invalid_argument("missing row arrays");
        if (!(tol >= 0.0)) throw std::invalid_argument("tol must be >= 0");
        const int n = lp->ncol, m = lp->nrow, nt = n + m;
        std::vector<double> act(static_cast<size_t>(m), 0.0);
        for (int j = 0; j < n; ++j)
            for (int64_t p = lp->col_start[j]; p < lp->col_start[j + 1]; ++p)
                act[lp->row_index[p]] += lp->value[p] * x[j];
        const double inf = std::numeric_limits<double>::infinity();
        std::vector<double> score(static_cast<size_t>(nt), -1.0);   // > = 0: basic candidate
        for (int v = 0; v < nt; ++v) {
            const double val = v < n ? x[v] : act[v - n];
            const double lo = v < n ? lp->collo[v] : lp->rowlo[v - n];
            const double up = v < n ? lp->colup[v] : lp->rowup[v - n];
            const bool lf = finite_bound(lo), uf = finite_bound(up);
            if (lf && uf && lo == up) {
                status[v] = NIR_VS_FIXED;
                continue;
            }
            const double dl = lf ? val - lo : inf, du = uf ? up - val : inf;
            const bool at_lo = lf && dl <= tol * (1.0 + std::fabs(lo));
            const bool at_up = uf && du <= tol * (1.0 + std::fabs(up));
            if (at_lo && (!at_up || dl <= du)) {
                status[v] = NIR_VS_AT_LOWER;
            } else if (at_up) {
                status[v] = NIR_VS_AT_UPPER;
            } else {
                status[v] = NIR_VS_BASIC;
                score[v] = (!lf && !uf) ? inf : std::min(dl, du) / (1.0 + std::fabs(val));
            }
        }
        // Nonbasic placement of a variable that loses (or never had) basis.
        auto nonbasic = [&](int v) {
            const double val = v < n ? x[v] : act[v - n];
            const double lo = v < n ? lp->collo[v] : lp->rowlo[v - n];
            const double up = v < n ? lp->colup[v] : lp->rowup[v - n];
            const bool lf = finite_bound(lo), uf = finite_bound(up);
            if (lf && uf) status[v] = (std::fabs(val - lo) <= std::fabs(up - val)) ? NIR_VS_AT_LOWER : NIR_VS_AT_UPPER;
            else if (lf) status[v] = NIR_VS_AT_LOWER;
            else if (uf) status[v] = NIR_VS_AT_UPPER;
            else status[v] = NIR_VS_FREE_ZERO;
        };
        std::vector<int> cand;
        for (int v = 0; v < nt; ++v)
            if (score[v] >= 0.0) cand.push_back(v);
        std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return score[a] > score[b]; });
        std::vector<char> basic(static_cast<size_t>(nt), 0);
        int count = 0;
        for (int v : cand) {
            if (count < m) {
                basic[v] = 1;
                ++count;
            } else {
                nonbasic(v);
            }
        }
        if (count < m) {
            // Logicals of the rows with the most slack (relative), then any.
            std::vector<int> rows(static_cast<size_t>(m));
            std::iota(rows.begin(), rows.end(), 0);
            auto slack = [&](int i) {
                const double lo = lp->rowlo[i], up = lp->rowup[i], a = act[i];
                const double dl = finite_bound(lo) ? a - lo : inf, du = finite_bound(up) ? up - a : inf;
                return std::min(dl, du) / (1.0 + std::fabs(a));
            };
            std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) { return slack(a) > slack(b); });
            for (int i : rows) {
                if (count >= m) break;
                if (!basic[n + i]) {
                    basic[n + i] = 1;
                    status[n + i] = NIR_VS_BASIC;
                    ++count;
                }
            }
        }
        int k = 0;
        for (int v = 0; v < nt; ++v)
            if (basic[v]) {
                status[v] = NIR_VS_BASIC;
                basic_index[k++] = v;
            }
        return 0;
    } catch (const std::exception& ex) {
        nir::record_last_error(ex.what());
        return -1;
    }
}

}  // extern "C"
