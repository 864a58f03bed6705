/* ORACLE - test infrastructure only (see oracle.h). Plain C99, compiled with
 * -ffp-contract=off. Each function cites the reference lines it restates. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* types.hpp:49-51 */
static int fin(double b) { return isfinite(b) && fabs(b) < 1e20; }
/* std::max / std::min as <algorithm> defines them */
static double smax(double a, double b) { return (a < b) ? b : a; }
static double smin(double a, double b) { return (b < a) ? b : a; }

/* csc_matrix.hpp:99-108: zero-fill, then scatter column by column (ascending
 * j), skipping columns with x_j == 0. */
void or_spmv(const or_lp* lp, const double* x, double* y) {
    for (int32_t i = 0; i < lp->m; ++i) y[i] = 0.0;
    for (int32_t j = 0; j < lp->n; ++j) {
        const double xj = x[j];
        if (xj == 0.0) continue;
        for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) y[lp->index[p]] += lp->value[p] * xj;
    }
}

/* csc_matrix.hpp:111-120: per-column sequential dot from 0.0. */
void or_spmv_t(const or_lp* lp, const double* y, double* g) {
    for (int32_t j = 0; j < lp->n; ++j) {
        double s = 0.0;
        for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) s += lp->value[p] * y[lp->index[p]];
        g[j] = s;
    }
}

/* scaling.hpp:14-19 */
static double pow2_round(double v) {
    if (v <= 0.0 || !isfinite(v)) return 1.0;
    int e = (int)round(log2(v));
    if (e < -30) e = -30;
    if (e > 30) e = 30;
    return ldexp(1.0, e);
}

/* scaling.hpp:56-127 (continuous columns only; the LPs here have no integers). */
int or_scale_geometric(const or_lp* lp, double* Rout, double* Cout) {
    const int32_t m = lp->m, n = lp->n;
    const int64_t nnz = n > 0 ? lp->start[n] : 0;
    for (int32_t i = 0; i < m; ++i) Rout[i] = 1.0;
    for (int32_t j = 0; j < n; ++j) Cout[j] = 1.0;
    if (nnz == 0) return 0;
    double hi = 0.0, lo = INFINITY;
    for (int64_t p = 0; p < nnz; ++p) {
        const double a = fabs(lp->value[p]);
        if (a > 0.0) {
            hi = smax(hi, a);
            lo = smin(lo, a);
        }
    }
    if (lo >= INFINITY || hi / lo <= 10.0) return 0;
    const int passes = (hi / lo < 1e3) ? 2 : 6;
    double* R = malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* C = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* rmin = malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* rmax = malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    for (int32_t i = 0; i < m; ++i) R[i] = 1.0;
    for (int32_t j = 0; j < n; ++j) C[j] = 1.0;
    for (int pass = 0; pass < passes; ++pass) {
        double change = 1.0;
        for (int32_t i = 0; i < m; ++i) {
            rmin[i] = INFINITY;
            rmax[i] = 0.0;
        }
        for (int32_t j = 0; j < n; ++j)
            for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) {
                const int32_t i = lp->index[p];
                const double v = fabs(lp->value[p]) * R[i] * C[j];
                if (v > 0.0) {
                    rmin[i] = smin(rmin[i], v);
                    rmax[i] = smax(rmax[i], v);
                }
            }
        for (int32_t i = 0; i < m; ++i)
            if (rmax[i] > 0.0 && rmin[i] < INFINITY) {
                const double f = 1.0 / sqrt(rmin[i] * rmax[i]);
                change = smax(change, smax(f, 1.0 / f));
                R[i] *= f;
            }
        for (int32_t j = 0; j < n; ++j) {
            double cmin = INFINITY, cmax = 0.0;
            for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) {
                const double v = fabs(lp->value[p]) * R[lp->index[p]] * C[j];
                if (v > 0.0) {
                    cmin = smin(cmin, v);
                    cmax = smax(cmax, v);
                }
            }
            if (cmax > 0.0 && cmin < INFINITY) {
                const double f = 1.0 / sqrt(cmin * cmax);
                change = smax(change, smax(f, 1.0 / f));
                C[j] *= f;
            }
        }
        if (change < 1.05) break;
    }
    for (int32_t i = 0; i < m; ++i) Rout[i] = pow2_round(R[i]);
    for (int32_t j = 0; j < n; ++j) Cout[j] = pow2_round(C[j]);
    free(R);
    free(C);
    free(rmin);
    free(rmax);
    return 1;
}

/* pdhg_solver.hpp:25-107 */
void or_pdhg_reference(const or_lp* lp, int64_t K, int check, const double* x0, const double* y0, double* x,
                       double* y, double* tau_out, int64_t* iters_out, double* viol_out) {
    const int32_t m = lp->m, n = lp->n;
    double* u = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* v = malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* g = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* xbar = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* h = malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));

    /* :32-45 power iteration */
    for (int32_t j = 0; j < n; ++j) u[j] = 1.0;
    double eta = 1.0;
    for (int p = 0; p < 12; ++p) {
        or_spmv(lp, u, v);
        or_spmv_t(lp, v, u);
        double n2 = 0.0;
        for (int32_t j = 0; j < n; ++j) n2 += u[j] * u[j];
        eta = sqrt(smax(1e-12, n2));
        for (int32_t j = 0; j < n; ++j) u[j] /= eta;
    }
    const double smx = sqrt(eta);
    const double tau = 0.85 / smax(1e-6, smx);
    const double sigma = 0.85 / smax(1e-6, smx);

    /* :47-57 start point */
    for (int32_t j = 0; j < n; ++j) {
        const double l = lp->l[j], up = lp->u[j];
        double s = 0.0;
        if (fin(l) && fin(up)) s = 0.5 * (l + up);
        else if (fin(l)) s = smax(0.0, l);
        else if (fin(up)) s = smin(0.0, up);
        x[j] = x0 ? x0[j] : s;
    }
    for (int32_t i = 0; i < m; ++i) y[i] = y0 ? y0[i] : 0.0;

    int64_t iters = 0;
    double V = 0.0;
    for (int64_t k = 0; k < K; ++k) {
        iters = k + 1;
        /* :70-78 primal step */
        or_spmv_t(lp, y, g);
        for (int32_t j = 0; j < n; ++j) {
            double cand = x[j] - tau * (lp->c[j] - g[j]);
            if (fin(lp->l[j])) cand = smax(lp->l[j], cand);
            if (fin(lp->u[j])) cand = smin(lp->u[j], cand);
            xbar[j] = 2.0 * cand - x[j];
            x[j] = cand;
        }
        /* :82-95 dual step and max primal violation */
        or_spmv(lp, xbar, h);
        V = 0.0;
        for (int32_t i = 0; i < m; ++i) {
            double r = h[i] - y[i] / sigma;
            if (fin(lp->rl[i])) r = smax(lp->rl[i], r);
            if (fin(lp->ru[i])) r = smin(lp->ru[i], r);
            y[i] += sigma * (r - h[i]);
            double viol = 0.0;
            if (fin(lp->rl[i])) viol = smax(viol, lp->rl[i] - h[i]);
            if (fin(lp->ru[i])) viol = smax(viol, h[i] - lp->ru[i]);
            V = smax(V, viol);
        }
        /* :97-100 stop rule */
        if (check && k > 20 && (k & 31) == 0 && V <= 1e-6) break;
    }
    if (tau_out) *tau_out = tau;
    if (iters_out) *iters_out = iters;
    if (viol_out) *viol_out = V;
    free(u);
    free(v);
    free(g);
    free(xbar);
    free(h);
}

/* ------------------------------------------------------------------ dd */
/* dd_real.hpp:8-24 error-free transformations; :58-82 accumulation. */
typedef struct {
    double hi, lo;
} dd;

static void dd_add(dd* a, double bh, double bl) {
    double s1 = a->hi + bh, bb = s1 - a->hi;
    double s2 = (a->hi - (s1 - bb)) + (bh - bb);
    double t1 = a->lo + bl, cc = t1 - a->lo;
    double t2 = (a->lo - (t1 - cc)) + (bl - cc);
    s2 += t1;
    double q = s1 + s2;
    s2 = s2 - (q - s1);
    s1 = q;
    s2 += t2;
    a->hi = s1 + s2;
    a->lo = s2 - (a->hi - s1);
}

static void dd_add_prod(dd* a, double x, double y) {
    const double p = x * y;
    dd_add(a, p, fma(x, y, -p));
}

/* sol_io.hpp:25-106 for continuous models (model.hpp:73-98 objective). */
int or_dd_verify(const or_lp* lp, const double* x, double claimed_obj, double tol, double* max_row_viol,
                 double* max_bound_viol, double* computed_obj) {
    const int32_t m = lp->m, n = lp->n;
    double bv = 0.0, rv = 0.0;
    for (int32_t j = 0; j < n; ++j) {
        if (!isfinite(x[j])) return 0;
        if (fin(lp->l[j])) bv = smax(bv, smax(0.0, lp->l[j] - x[j]) / (1.0 + fabs(lp->l[j])));
        if (fin(lp->u[j])) bv = smax(bv, smax(0.0, x[j] - lp->u[j]) / (1.0 + fabs(lp->u[j])));
    }
    dd* act = calloc((size_t)(m > 0 ? m : 1), sizeof(dd));
    for (int32_t j = 0; j < n; ++j) {
        if (x[j] == 0.0) continue;
        for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) dd_add_prod(&act[lp->index[p]], lp->value[p], x[j]);
    }
    for (int32_t i = 0; i < m; ++i) {
        const double r = act[i].hi + act[i].lo;
        if (fin(lp->rl[i])) rv = smax(rv, smax(0.0, lp->rl[i] - r) / (1.0 + fabs(lp->rl[i])));
        if (fin(lp->ru[i])) rv = smax(rv, smax(0.0, r - lp->ru[i]) / (1.0 + fabs(lp->ru[i])));
    }
    free(act);
    dd obj = {lp->offset, 0.0};
    for (int32_t j = 0; j < n; ++j)
        if (lp->c[j] != 0.0 && x[j] != 0.0) dd_add_prod(&obj, lp->c[j], x[j]);
    const double cobj = obj.hi + obj.lo;
    double rel = 0.0;
    if (isfinite(claimed_obj)) rel = fabs(cobj - claimed_obj) / (1.0 + fabs(cobj));
    if (max_row_viol) *max_row_viol = rv;
    if (max_bound_viol) *max_bound_viol = bv;
    if (computed_obj) *computed_obj = cobj;
    return !(bv > tol || rv > tol || rel > 1e-5);
}

/* ------------------------------------------------------------------ KKT */
/* Relative KKT error of the PDLP family (build-defined, evaluated in the LP's
 * own units; sums in double-double):
 *   primal  ||A x - clamp(A x, rl, ru)||_2 / (1 + ||b||_2),
 *           b_i = the finite row bound of largest magnitude (0 if none)
 *   dual    ||(c - A^T y) - proj(c - A^T y)||_2 and the sign violations of y,
 *           over (1 + ||c||_2); proj keeps the reduced-cost signs the column
 *           bounds allow (free sign when both are finite)
 *   gap     |P - D| / (1 + |P| + |D|), P = c^T x + off,
 *           D = off + sum_i (rl_i y_i^+ - ru_i y_i^-) + sum_j (l_j lam_j^+ - u_j lam_j^-)
 * out = {rel_primal, rel_dual, rel_gap, P, D}. */
void or_kkt(const or_lp* lp, const double* x, const double* y, double* out) {
    const int32_t m = lp->m, n = lp->n;
    double* ax = malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* aty = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    /* exact-as-possible products for the residuals */
    for (int32_t i = 0; i < m; ++i) ax[i] = 0.0;
    {
        dd* acc = calloc((size_t)(m > 0 ? m : 1), sizeof(dd));
        for (int32_t j = 0; j < n; ++j)
            for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) dd_add_prod(&acc[lp->index[p]], lp->value[p], x[j]);
        for (int32_t i = 0; i < m; ++i) ax[i] = acc[i].hi + acc[i].lo;
        free(acc);
    }
    for (int32_t j = 0; j < n; ++j) {
        dd s = {0.0, 0.0};
        for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) dd_add_prod(&s, lp->value[p], y[lp->index[p]]);
        aty[j] = s.hi + s.lo;
    }
    dd rp = {0, 0}, rd = {0, 0}, bn = {0, 0}, cn = {0, 0}, P = {lp->offset, 0}, D = {lp->offset, 0};
    for (int32_t i = 0; i < m; ++i) {
        const int lf = fin(lp->rl[i]), uf = fin(lp->ru[i]);
        double proj = ax[i];
        if (lf) proj = smax(lp->rl[i], proj);
        if (uf) proj = smin(lp->ru[i], proj);
        const double r = ax[i] - proj;
        dd_add_prod(&rp, r, r);
        double b = 0.0;
        if (lf) b = fabs(lp->rl[i]);
        if (uf) b = smax(b, fabs(lp->ru[i]));
        dd_add_prod(&bn, b, b);
        /* sign-feasible part of y and its violation */
        double yp = y[i];
        if (!lf) yp = smin(yp, 0.0);
        if (!uf) yp = smax(yp, 0.0);
        const double yv = y[i] - yp;
        dd_add_prod(&rd, yv, yv);
        if (lf && yp > 0.0) dd_add_prod(&D, lp->rl[i], yp);
        if (uf && yp < 0.0) dd_add_prod(&D, lp->ru[i], yp);
    }
    for (int32_t j = 0; j < n; ++j) {
        const int lf = fin(lp->l[j]), uf = fin(lp->u[j]);
        const double lam = lp->c[j] - aty[j];
        double lp_ = lam;
        if (!lf) lp_ = smin(lp_, 0.0);
        if (!uf) lp_ = smax(lp_, 0.0);
        const double r = lam - lp_;
        dd_add_prod(&rd, r, r);
        dd_add_prod(&cn, lp->c[j], lp->c[j]);
        if (x[j] != 0.0) dd_add_prod(&P, lp->c[j], x[j]);
        if (lf && lp_ > 0.0) dd_add_prod(&D, lp->l[j], lp_);
        if (uf && lp_ < 0.0) dd_add_prod(&D, lp->u[j], lp_);
    }
    const double pv = P.hi + P.lo, dv = D.hi + D.lo;
    out[0] = sqrt(rp.hi + rp.lo) / (1.0 + sqrt(bn.hi + bn.lo));
    out[1] = sqrt(rd.hi + rd.lo) / (1.0 + sqrt(cn.hi + cn.lo));
    out[2] = fabs(pv - dv) / (1.0 + fabs(pv) + fabs(dv));
    out[3] = pv;
    out[4] = dv;
    free(ax);
    free(aty);
}
