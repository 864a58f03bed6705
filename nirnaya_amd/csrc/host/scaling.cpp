#include <algorithm>
#include <cmath>

#include "prep.hpp"

namespace nir {

double Scaling::to_pow2(double v) {
    if (v <= 0.0 || !std::isfinite(v)) return 1.0;
    int e = static_cast<int>(std::round(std::log2(v)));
    e = std::max(-30, std::min(30, e));
    return std::ldexp(1.0, e);
}

void Scaling::unscale_primal(const std::vector<double>& xs, std::vector<double>& x) const {
    x.resize(xs.size());
    for (size_t j = 0; j < xs.size(); ++j) x[j] = active ? xs[j] * col[j] : xs[j];
}

void Scaling::unscale_duals(const std::vector<double>& ys, std::vector<double>& y) const {
    y.resize(ys.size());
    for (size_t i = 0; i < ys.size(); ++i) y[i] = active ? ys[i] * row[i] : ys[i];
}

Model scale_geometric(const Model& orig, bool enable, Scaling& info) {
    const Idx m = orig.nrow, n = orig.ncol;
    info.row.assign(static_cast<size_t>(m), 1.0);
    info.col.assign(static_cast<size_t>(n), 1.0);
    info.active = false;
    if (!enable || orig.A.nnz() == 0) return orig;

    double hi = 0.0, lo = kInf;
    for (double v : orig.A.value) {
        const double a = std::fabs(v);
        if (a > 0.0) {
            hi = amax(hi, a);
            lo = amin(lo, a);
        }
    }
    if (lo >= kInf || hi / lo <= 10.0) return orig;

    info.active = true;
    const int passes = (hi / lo < 1e3) ? 2 : 6;
    std::vector<double> R(static_cast<size_t>(m), 1.0), C(static_cast<size_t>(n), 1.0);
    const CscMatrix& A = orig.A;

    for (int pass = 0; pass < passes; ++pass) {
        double worst = 1.0;   // largest factor (or inverse) applied this pass
        // Rows: 1/sqrt(min * max) of |a_ij| R_i C_j over the row.
        std::vector<double> rmin(static_cast<size_t>(m), kInf), rmax(static_cast<size_t>(m), 0.0);
        for (Idx j = 0; j < n; ++j) {
            for (Off p = A.start[j]; p < A.start[j + 1]; ++p) {
                const Idx i = A.index[p];
                const double v = std::fabs(A.value[p]) * R[i] * C[j];
                if (v > 0.0) {
                    rmin[i] = amin(rmin[i], v);
                    rmax[i] = amax(rmax[i], v);
                }
            }
        }
        for (Idx i = 0; i < m; ++i) {
            if (rmax[i] > 0.0 && rmin[i] < kInf) {
                const double f = 1.0 / std::sqrt(rmin[i] * rmax[i]);
                worst = amax(worst, amax(f, 1.0 / f));
                R[i] *= f;
            }
        }
        // Columns (continuous only), using the freshly updated row factors.
        for (Idx j = 0; j < n; ++j) {
            if (orig.vartype[j] != VarType::Continuous) continue;
            double cmin = kInf, cmax = 0.0;
            for (Off p = A.start[j]; p < A.start[j + 1]; ++p) {
                const double v = std::fabs(A.value[p]) * R[A.index[p]] * C[j];
                if (v > 0.0) {
                    cmin = amin(cmin, v);
                    cmax = amax(cmax, v);
                }
            }
            if (cmax > 0.0 && cmin < kInf) {
                const double f = 1.0 / std::sqrt(cmin * cmax);
                worst = amax(worst, amax(f, 1.0 / f));
                C[j] *= f;
            }
        }
        if (worst < 1.05) break;
    }

    for (Idx i = 0; i < m; ++i) info.row[i] = Scaling::to_pow2(R[i]);
    for (Idx j = 0; j < n; ++j)
        info.col[j] = orig.vartype[j] == VarType::Continuous ? Scaling::to_pow2(C[j]) : 1.0;

    Model s = orig;
    for (Idx j = 0; j < n; ++j) {
        const double cj = info.col[j];
        s.cost[j] = orig.cost[j] * cj;
        if (finite_bound(orig.collo[j])) s.collo[j] = orig.collo[j] / cj;
        if (finite_bound(orig.colup[j])) s.colup[j] = orig.colup[j] / cj;
        for (Off p = s.A.start[j]; p < s.A.start[j + 1]; ++p)
            s.A.value[p] = orig.A.value[p] * info.row[s.A.index[p]] * cj;
    }
    for (Idx i = 0; i < m; ++i) {
        const double ri = info.row[i];
        if (finite_bound(orig.rowlo[i])) s.rowlo[i] = orig.rowlo[i] * ri;
        if (finite_bound(orig.rowup[i])) s.rowup[i] = orig.rowup[i] * ri;
    }
    if (s.Q.nnz() > 0) {
        for (Idx j = 0; j < n; ++j)
            for (Off p = s.Q.start[j]; p < s.Q.start[j + 1]; ++p)
                s.Q.value[p] = orig.Q.value[p] * info.col[s.Q.index[p]] * info.col[j];
    }
    return s;
}

}  // namespace nir
