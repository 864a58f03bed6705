#include <cmath>

#include "prep.hpp"

namespace nir {

void Presolved::postsolve(const std::vector<double>& x_red, std::vector<double>& x_orig) const {
    x_orig = fixed_x;
    for (size_t k = 0; k < col_map.size() && k < x_red.size(); ++k) x_orig[col_map[k]] = x_red[k];
}

namespace {

// Working state of one presolve run. Rules mutate bounds, offsets and the
// deletion masks; the matrix values never change.
struct Reducer {
    Model w;
    CsrMatrix rows;
    std::vector<char> row_gone, col_gone;
    Presolved& out;

    Reducer(const Model& orig, Presolved& res)
        : w(orig), rows(CsrMatrix::transpose_of(orig.A)),
          row_gone(static_cast<size_t>(orig.nrow), 0), col_gone(static_cast<size_t>(orig.ncol), 0),
          out(res) {}

    // Remove column j at value v: fold its cost into the offset and its
    // activity into the finite bounds of the rows still present.
    void fix(Idx j, double v) {
        if (col_gone[j]) return;
        col_gone[j] = 1;
        out.fixed_x[j] = v;
        w.obj_offset += w.cost[j] * v;
        for (Off p = w.A.start[j]; p < w.A.start[j + 1]; ++p) {
            const Idx i = w.A.index[p];
            if (row_gone[i]) continue;
            const double a = w.A.value[p];
            if (finite_bound(w.rowlo[i])) w.rowlo[i] -= a * v;
            if (finite_bound(w.rowup[i])) w.rowup[i] -= a * v;
        }
        ++out.cols_removed;
    }

    bool verdict(Status s) {
        out.status = s;
        out.solved = true;
        return true;
    }

    // R4: columns whose bounds coincide within 1e-11.
    bool fixed_columns() {
        bool changed = false;
        for (Idx j = 0; j < w.ncol; ++j) {
            if (col_gone[j]) continue;
            if (finite_bound(w.collo[j]) && finite_bound(w.colup[j]) &&
                std::fabs(w.colup[j] - w.collo[j]) <= 1e-11) {
                fix(j, w.collo[j]);
                changed = true;
            }
        }
        return changed;
    }

    // R1 / R3: empty rows (feasibility check) and singleton rows (turned into
    // column bounds). Returns true when a verdict ends presolve.
    bool short_rows(bool& changed) {
        for (Idx i = 0; i < w.nrow; ++i) {
            if (row_gone[i]) continue;
            Idx count = 0, only = -1;
            double a_only = 0.0;
            for (Off p = rows.start[i]; p < rows.start[i + 1]; ++p) {
                const Idx j = rows.index[p];
                if (!col_gone[j] && std::fabs(rows.value[p]) > 1e-13) {
                    ++count;
                    only = j;
                    a_only = rows.value[p];
                }
            }
            if (count == 0) {
                if ((finite_bound(w.rowlo[i]) && w.rowlo[i] > 1e-7) ||
                    (finite_bound(w.rowup[i]) && w.rowup[i] < -1e-7))
                    return verdict(Status::Infeasible);
                row_gone[i] = 1;
                ++out.rows_removed;
                changed = true;
            } else if (count == 1) {
                double lo = -kInf, up = kInf;
                if (a_only > 0.0) {
                    if (finite_bound(w.rowlo[i])) lo = w.rowlo[i] / a_only;
                    if (finite_bound(w.rowup[i])) up = w.rowup[i] / a_only;
                } else {
                    if (finite_bound(w.rowup[i])) lo = w.rowup[i] / a_only;
                    if (finite_bound(w.rowlo[i])) up = w.rowlo[i] / a_only;
                }
                if (w.vartype[only] != VarType::Continuous) {
                    if (finite_bound(lo)) lo = std::ceil(lo - 1e-7);
                    if (finite_bound(up)) up = std::floor(up + 1e-7);
                }
                if (lo > w.collo[only] + 1e-9) {
                    w.collo[only] = lo;
                    ++out.tightenings;
                }
                if (up < w.colup[only] - 1e-9) {
                    w.colup[only] = up;
                    ++out.tightenings;
                }
                if (w.collo[only] > w.colup[only] + 1e-7) return verdict(Status::Infeasible);
                if (w.collo[only] > w.colup[only]) w.collo[only] = w.colup[only];
                row_gone[i] = 1;
                ++out.rows_removed;
                changed = true;
            }
        }
        return false;
    }

    // Integer bound tightening from a finite activity bound (R7).
    // `slack` is row-bound minus extreme activity; `upper_side` selects the
    // <= side (min activity) or the >= side (max activity).
    void tighten_integers(Idx i, double slack, bool upper_side, bool& changed) {
        for (Off p = rows.start[i]; p < rows.start[i + 1]; ++p) {
            const Idx j = rows.index[p];
            if (col_gone[j] || w.vartype[j] == VarType::Continuous) continue;
            const double a = rows.value[p];
            if (upper_side) {
                if (a > 1e-9 && finite_bound(w.collo[j])) {
                    const double nu = std::floor(w.collo[j] + (slack / a) + 1e-6);
                    if (nu < w.colup[j] - 0.5) { w.colup[j] = nu; ++out.tightenings; changed = true; }
                } else if (a < -1e-9 && finite_bound(w.colup[j])) {
                    const double nl = std::ceil(w.colup[j] + (slack / a) - 1e-6);
                    if (nl > w.collo[j] + 0.5) { w.collo[j] = nl; ++out.tightenings; changed = true; }
                }
            } else {
                if (a > 1e-9 && finite_bound(w.colup[j])) {
                    const double nl = std::ceil(w.colup[j] - (slack / a) - 1e-6);
                    if (nl > w.collo[j] + 0.5) { w.collo[j] = nl; ++out.tightenings; changed = true; }
                } else if (a < -1e-9 && finite_bound(w.collo[j])) {
                    const double nu = std::floor(w.collo[j] - (slack / a) + 1e-6);
                    if (nu < w.colup[j] - 0.5) { w.colup[j] = nu; ++out.tightenings; changed = true; }
                }
            }
        }
    }

    // R5 / R6 / R14 / R7 on activity bounds. Returns true on a verdict.
    bool activity_rules(bool& changed) {
        for (Idx i = 0; i < w.nrow; ++i) {
            if (row_gone[i]) continue;
            double amin_act = 0.0, amax_act = 0.0;
            int amin_inf = 0, amax_inf = 0;
            for (Off p = rows.start[i]; p < rows.start[i + 1]; ++p) {
                const Idx j = rows.index[p];
                if (col_gone[j]) continue;
                const double a = rows.value[p];
                const double toward_min = a > 0.0 ? w.collo[j] : w.colup[j];
                const double toward_max = a > 0.0 ? w.colup[j] : w.collo[j];
                if (a == 0.0) continue;
                if (finite_bound(toward_min)) amin_act += a * toward_min; else ++amin_inf;
                if (finite_bound(toward_max)) amax_act += a * toward_max; else ++amax_inf;
            }
            const bool lo_fin = finite_bound(w.rowlo[i]), up_fin = finite_bound(w.rowup[i]);
            // R6: activity range misses the row range.
            if ((amin_inf == 0 && up_fin && amin_act > w.rowup[i] + 1e-7) ||
                (amax_inf == 0 && lo_fin && amax_act < w.rowlo[i] - 1e-7))
                return verdict(Status::Infeasible);
            // R5: row implied by column bounds.
            const bool lo_implied = !lo_fin || (amin_inf == 0 && amin_act >= w.rowlo[i] - 1e-9);
            const bool up_implied = !up_fin || (amax_inf == 0 && amax_act <= w.rowup[i] + 1e-9);
            if (lo_implied && up_implied) {
                row_gone[i] = 1;
                ++out.rows_removed;
                changed = true;
                continue;
            }
            // R14: forcing rows pin every column at the bound that attains it.
            if (amin_inf == 0 && up_fin && std::fabs(amin_act - w.rowup[i]) <= 1e-9) {
                for (Off p = rows.start[i]; p < rows.start[i + 1]; ++p) {
                    const Idx j = rows.index[p];
                    if (col_gone[j]) continue;
                    fix(j, rows.value[p] > 0.0 ? w.collo[j] : w.colup[j]);
                }
                row_gone[i] = 1;
                ++out.rows_removed;
                changed = true;
                continue;
            }
            if (amax_inf == 0 && lo_fin && std::fabs(amax_act - w.rowlo[i]) <= 1e-9) {
                for (Off p = rows.start[i]; p < rows.start[i + 1]; ++p) {
                    const Idx j = rows.index[p];
                    if (col_gone[j]) continue;
                    fix(j, rows.value[p] > 0.0 ? w.colup[j] : w.collo[j]);
                }
                row_gone[i] = 1;
                ++out.rows_removed;
                changed = true;
                continue;
            }
            if (amin_inf == 0 && up_fin) tighten_integers(i, w.rowup[i] - amin_act, true, changed);
            if (amax_inf == 0 && lo_fin) tighten_integers(i, amax_act - w.rowlo[i], false, changed);
        }
        return false;
    }

    // R2: columns with no remaining row go to their cost-optimal bound.
    bool empty_columns(bool& changed) {
        for (Idx j = 0; j < w.ncol; ++j) {
            if (col_gone[j]) continue;
            bool touches = false;
            for (Off p = w.A.start[j]; p < w.A.start[j + 1]; ++p)
                if (!row_gone[w.A.index[p]]) { touches = true; break; }
            if (touches) continue;
            const double c = w.cost[j];
            if (c > 1e-12) {
                if (!finite_bound(w.collo[j])) return verdict(Status::Unbounded);
                fix(j, w.collo[j]);
            } else if (c < -1e-12) {
                if (!finite_bound(w.colup[j])) return verdict(Status::Unbounded);
                fix(j, w.colup[j]);
            } else {
                double v = 0.0;
                if (finite_bound(w.collo[j]) && v < w.collo[j]) v = w.collo[j];
                if (finite_bound(w.colup[j]) && v > w.colup[j]) v = w.colup[j];
                fix(j, v);
            }
            changed = true;
        }
        return false;
    }
};

}  // namespace

Presolved presolve(const Model& orig, int level, int max_rounds) {
    Presolved res;
    res.fixed_x.assign(static_cast<size_t>(orig.ncol), 0.0);
    if (level <= 0 || orig.is_qp()) {
        res.reduced = orig;
        for (Idx j = 0; j < orig.ncol; ++j) res.col_map.push_back(j);
        for (Idx i = 0; i < orig.nrow; ++i) res.row_map.push_back(i);
        return res;
    }

    Reducer r(orig, res);
    for (int round = 0; round < max_rounds; ++round) {
        bool changed = r.fixed_columns();
        if (r.short_rows(changed)) return res;
        if (r.activity_rules(changed)) return res;
        if (r.empty_columns(changed)) return res;
        if (!changed) break;
    }

    Model& red = res.reduced;
    red = Model();
    red.name = orig.name + "_pre";
    red.sense = orig.sense;
    red.obj_offset = r.w.obj_offset;
    std::vector<Idx> new_col(static_cast<size_t>(orig.ncol), -1), new_row(static_cast<size_t>(orig.nrow), -1);
    for (Idx j = 0; j < orig.ncol; ++j) {
        if (r.col_gone[j]) continue;
        new_col[j] = red.ncol++;
        res.col_map.push_back(j);
        red.cost.push_back(r.w.cost[j]);
        red.collo.push_back(r.w.collo[j]);
        red.colup.push_back(r.w.colup[j]);
        red.vartype.push_back(r.w.vartype[j]);
        red.col_names.push_back(r.w.col_names[j]);
    }
    for (Idx i = 0; i < orig.nrow; ++i) {
        if (r.row_gone[i]) continue;
        new_row[i] = red.nrow++;
        res.row_map.push_back(i);
        red.rowlo.push_back(r.w.rowlo[i]);
        red.rowup.push_back(r.w.rowup[i]);
        red.row_names.push_back(r.w.row_names[i]);
    }
    std::vector<Entry> kept;
    for (Idx j = 0; j < orig.ncol; ++j) {
        if (r.col_gone[j]) continue;
        for (Off p = r.w.A.start[j]; p < r.w.A.start[j + 1]; ++p) {
            const Idx i = r.w.A.index[p];
            if (!r.row_gone[i]) kept.push_back({new_row[i], new_col[j], r.w.A.value[p]});
        }
    }
    red.A = CscMatrix::assemble(red.nrow, red.ncol, kept);
    red.Q = CscMatrix(red.ncol, red.ncol);
    if (red.ncol == 0 && red.nrow == 0) {
        res.solved = true;
        res.status = Status::Optimal;
    }
    return res;
}

}  // namespace nir

