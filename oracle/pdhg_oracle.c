/* ORACLE - test infrastructure only (see oracle.h). Plain C99, compiled with
 * -ffp-contract=off. Each function cites the reference lines it restates. */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
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
/* The loop of pdhg_solver.hpp:59-101 from (x, y): K iterations (fewer when
 * check != 0 and the stop rule fires). Returns the last max row violation. */
static double ref_iterations(const or_lp* lp, int64_t K, int check, double tau, double sigma, double* x, double* y,
                             double* g, double* xbar, double* h, int64_t* iters_out) {
    const int32_t m = lp->m, n = lp->n;
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
    *iters_out = iters;
    return V;
}

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
    const double V = ref_iterations(lp, K, check, tau, sigma, x, y, g, xbar, h, &iters);
    if (tau_out) *tau_out = tau;
    if (iters_out) *iters_out = iters;
    if (viol_out) *viol_out = V;
    free(u);
    free(v);
    free(g);
    free(xbar);
    free(h);
}

void or_pdhg_steps(const or_lp* lp, int64_t K, double tau, double sigma, double* x, double* y) {
    const int32_t m = lp->m, n = lp->n;
    double* g = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* xbar = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* h = malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    int64_t iters = 0;
    (void)ref_iterations(lp, K, 0, tau, sigma, x, y, g, xbar, h, &iters);
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

/* ------------------------------------------------------------------ PDLP */
/* Restatement of the engine's "pdlp" mode (build-defined, no reference
 * counterpart): exponent-domain Ruiz (10 passes) + Pock-Chambolle scaling,
 * power iteration for eta = 0.998 / sigma_max, reflected Halpern PDHG with
 * KKT restarts every eval_freq iterations and primal weight
 * omega <- sqrt(omega ||dy|| / ||dx||). Sums of A products are sequential in
 * ascending index order from 0.0 (as the device's CSR-stream kernels);
 * norms and objectives are double-double. */

static double pow2_inv_sqrt(double v) {
    if (!(v > 0.0) || !isfinite(v)) return 1.0;
    int e2;
    (void)frexp(v, &e2);
    const int e = e2 - 1;
    const int k = (e + 1) >= 0 ? (e + 1) / 2 : -((-(e + 1) + 1) / 2);
    return ldexp(1.0, -k);
}

static dd dd_sum(dd a, dd b) {
    dd r = a;
    dd_add(&r, b.hi, b.lo);
    return r;
}

static dd ddp(double a, double b) {
    const double p = a * b;
    dd r = {p, fma(a, b, -p)};
    return r;
}

static double ddv(dd a) { return a.hi + a.lo; }

typedef struct {
    double omega, tau, sigma, eta, kkt_anchor, kkt_prev, rel_p, rel_d, rel_g, pobj, dobj, tol, bnorm_o, cnorm_o, offset;
    int64_t total, max_iters;
    int k_base, done, status, restart, restarts, eval_freq;
} pstate;

typedef struct {
    int32_t m, n;
    int32_t *rp, *ci;   /* CSR(A) */
    double *av;         /* CSR values (scaled) */
    const int64_t* cs;  /* CSC(A) = CSR(A^T) */
    const int32_t* ri;
    double* tv;         /* CSC values (scaled) */
} pmat;

/* The engine's pdlp-mode order of a sum over one row of L nonzeros
 * (engine/spmv.hpp, stream_rows with ORDER == 1, and the split-row kernels
 * of engine/pdlp.hip): for L <= 4096 the sequential sum from 0.0 in
 * ascending position; for L > 4096, terms are summed from 0.0 in chunks of
 * 64 consecutive positions, the chunk sums from 0.0 in groups of 32 chunks,
 * and the group sums from 0.0 into the row sum. The term at position p is
 * val[p] * v[idx[p]], or |val[p]| when v is NULL (the Pock-Chambolle row
 * measure). val/idx point at the row's first entry. */
#define OR_LONG_ROW 4096
#define OR_CHUNK 64
#define OR_GROUP 32
static double pdlp_term(const double* val, const int32_t* idx, const double* v, int64_t p) {
    return v ? val[p] * v[idx[p]] : fabs(val[p]);
}

double or_pdlp_row_sum(const double* val, const int32_t* idx, int64_t len, const double* v) {
    double s = 0.0;
    if (len <= OR_LONG_ROW) {
        for (int64_t p = 0; p < len; ++p) s += pdlp_term(val, idx, v, p);
        return s;
    }
    double group = 0.0;
    int in_group = 0;
    for (int64_t c0 = 0; c0 < len; c0 += OR_CHUNK) {
        const int64_t c1 = c0 + OR_CHUNK < len ? c0 + OR_CHUNK : len;
        double cs = 0.0;
        for (int64_t p = c0; p < c1; ++p) cs += pdlp_term(val, idx, v, p);
        group += cs;
        if (++in_group == OR_GROUP) {
            s += group;
            group = 0.0;
            in_group = 0;
        }
    }
    if (in_group > 0) s += group;
    return s;
}

static double row_dot(const pmat* M, int32_t i, const double* v) {
    return or_pdlp_row_sum(M->av + M->rp[i], M->ci + M->rp[i], (int64_t)M->rp[i + 1] - M->rp[i], v);
}

static double col_dot(const pmat* M, int32_t j, const double* v) {
    return or_pdlp_row_sum(M->tv + M->cs[j], M->ri + M->cs[j], M->cs[j + 1] - M->cs[j], v);
}

/* Evaluation at (xp, yp) with anchor (xa, ya); mode 1 = start point. */
static void pdlp_eval(const pmat* M, pstate* s, const double* c, const double* lo, const double* up, const double* rl,
                      const double* ru, const double* rt, const double* ct, const double* xp, const double* yp,
                      const double* xa, const double* ya, int mode) {
    dd rd2 = {0, 0}, rd2o = {0, 0}, Pc = {0, 0}, Dc = {0, 0}, dx2 = {0, 0};
    dd rp2 = {0, 0}, rp2o = {0, 0}, yv2 = {0, 0}, yv2o = {0, 0}, Dr = {0, 0}, dy2 = {0, 0};
    for (int32_t j = 0; j < M->n; ++j) {
        const double g = col_dot(M, j, yp);
        const double lam = c[j] - g;
        const int lf = lo[j] > -INFINITY, uf = up[j] < INFINITY;
        double lp = lam;
        if (!lf) lp = smin(lp, 0.0);
        if (!uf) lp = smax(lp, 0.0);
        const double rd = lam - lp;
        rd2 = dd_sum(rd2, ddp(rd, rd));
        const double rdo = rd / ct[j];
        rd2o = dd_sum(rd2o, ddp(rdo, rdo));
        Pc = dd_sum(Pc, ddp(c[j], xp[j]));
        if (lf && lp > 0.0) Dc = dd_sum(Dc, ddp(lo[j], lp));
        if (uf && lp < 0.0) Dc = dd_sum(Dc, ddp(up[j], lp));
        const double dx = xp[j] - xa[j];
        dx2 = dd_sum(dx2, ddp(dx, dx));
    }
    for (int32_t i = 0; i < M->m; ++i) {
        const double h = row_dot(M, i, xp);
        const int lf = rl[i] > -INFINITY, uf = ru[i] < INFINITY;
        double proj = smax(rl[i], h);
        proj = smin(ru[i], proj);
        const double r = h - proj;
        rp2 = dd_sum(rp2, ddp(r, r));
        const double ro = r / rt[i];
        rp2o = dd_sum(rp2o, ddp(ro, ro));
        double ys = yp[i];
        if (!lf) ys = smin(ys, 0.0);
        if (!uf) ys = smax(ys, 0.0);
        const double yv = yp[i] - ys;
        yv2 = dd_sum(yv2, ddp(yv, yv));
        const double yvo = yv * rt[i];
        yv2o = dd_sum(yv2o, ddp(yvo, yvo));
        if (lf && ys > 0.0) Dr = dd_sum(Dr, ddp(rl[i], ys));
        if (uf && ys < 0.0) Dr = dd_sum(Dr, ddp(ru[i], ys));
        const double dy = yp[i] - ya[i];
        dy2 = dd_sum(dy2, ddp(dy, dy));
    }
    dd off = {s->offset, 0.0};
    const double vrd2 = ddv(dd_sum(rd2, yv2)), vrd2o = ddv(dd_sum(rd2o, yv2o));
    const double vrp2 = ddv(rp2), vrp2o = ddv(rp2o);
    const double P = ddv(dd_sum(Pc, off)), D = ddv(dd_sum(dd_sum(Dc, Dr), off));
    const double vdx2 = ddv(dx2), vdy2 = ddv(dy2);
    const double gap = fabs(P - D);
    s->rel_p = sqrt(vrp2o) / (1.0 + s->bnorm_o);
    s->rel_d = sqrt(vrd2o) / (1.0 + s->cnorm_o);
    s->rel_g = gap / (1.0 + fabs(P) + fabs(D));
    s->pobj = P;
    s->dobj = D;
    const double worst = smax(s->rel_p, smax(s->rel_d, s->rel_g));
    s->restart = 0;
    if (!isfinite(vrd2) || !isfinite(vrp2) || !isfinite(P) || !isfinite(D) || !isfinite(vdx2) || !isfinite(vdy2)) {
        s->done = 1;
        s->status = 11;
        return;
    }
    const double kkt = sqrt(s->omega * vrp2 + vrd2 / s->omega + gap * gap);
    if (mode == 1) {
        s->kkt_anchor = kkt;
        s->kkt_prev = kkt;
        if (worst <= s->tol) {
            s->done = 1;
            s->status = 0;
        }
        return;
    }
    s->total += s->eval_freq;
    if (worst <= s->tol) {
        s->done = 1;
        s->status = 0;
        return;
    }
    if (s->total >= s->max_iters) {
        s->done = 1;
        s->status = 6;
        return;
    }
    const int k_new = s->k_base + s->eval_freq;
    const int restart = (kkt <= 0.2 * s->kkt_anchor) || (kkt <= 0.8 * s->kkt_anchor && kkt > s->kkt_prev) ||
                        ((double)k_new >= 0.36 * (double)s->total);
    s->kkt_prev = kkt;
    if (restart) {
        if (vdx2 > 1e-20 && vdy2 > 1e-20) {
            const double ratio = sqrt(vdy2) / sqrt(vdx2);
            s->omega = sqrt(s->omega * ratio);
            s->tau = s->eta / s->omega;
            s->sigma = s->eta * s->omega;
        }
        s->kkt_anchor = sqrt(s->omega * vrp2 + vrd2 / s->omega + gap * gap);
        s->kkt_prev = s->kkt_anchor;
        s->k_base = 0;
        s->restarts += 1;
        s->restart = 1;
    } else {
        s->k_base = k_new;
    }
}

/* Infeasibility certificate test of a candidate ray (dx, dy) in the scaled
 * space (build-defined, after the PDLP infeasibility-detection idea: on an
 * infeasible or unbounded LP the restarted iterates drift along a ray, so the
 * difference of consecutive restart points is tested as one).
 *   Dual ray dy (certifies primal infeasibility): lam = -A^T dy; violations
 *   are lam - proj(lam) for infinite column bounds and dy_i > 0 / dy_i < 0 for
 *   infinite rl_i / ru_i; the ray objective
 *   D = sum_i (rl_i dy_i [dy_i > 0] + ru_i dy_i [dy_i < 0])
 *     + sum_j (l_j lp_j [lp_j > 0] + u_j lp_j [lp_j < 0])
 *   must be positive.
 *   Primal ray dx (certifies dual infeasibility): violations are the parts of
 *   A dx and dx outside the recession cones of the rows and bounds; -c^T dx
 *   must be positive.
 * A dx and A^T dy are double-double dot products in ascending index order
 * (chunked as in or_pdlp_row_sum for lines of more than 4096 nonzeros) and
 * every sum is double-double, so an exact zero stays zero to ~1e-32. A ray
 * passes when its objective exceeds 1e-12 of the sum of its terms' magnitudes
 * and ||violations||_2 <= eps * objective. Returns 21 (dual ray passes), 22
 * (primal ray passes) or 0; ratios[0..1] = violation / objective (inf when
 * the objective test fails). */
static double dd_dot_line(const double* val, const int32_t* idx, int64_t len, const double* v) {
    dd s = {0, 0};
    if (len <= OR_LONG_ROW) {
        for (int64_t p = 0; p < len; ++p) dd_add_prod(&s, val[p], v[idx[p]]);
        return s.hi + s.lo;
    }
    /* Lines of more than 4096 nonzeros: the chunked order of or_pdlp_row_sum
     * with every level in double-double. */
    dd group = {0, 0};
    int in_group = 0;
    for (int64_t c0 = 0; c0 < len; c0 += OR_CHUNK) {
        const int64_t c1 = c0 + OR_CHUNK < len ? c0 + OR_CHUNK : len;
        dd cs = {0, 0};
        for (int64_t p = c0; p < c1; ++p) dd_add_prod(&cs, val[p], v[idx[p]]);
        dd_add(&group, cs.hi, cs.lo);
        if (++in_group == OR_GROUP) {
            dd_add(&s, group.hi, group.lo);
            group.hi = group.lo = 0.0;
            in_group = 0;
        }
    }
    if (in_group > 0) dd_add(&s, group.hi, group.lo);
    return s.hi + s.lo;
}

static double row_dot_dd(const pmat* M, int32_t i, const double* v) {
    return dd_dot_line(M->av + M->rp[i], M->ci + M->rp[i], (int64_t)M->rp[i + 1] - M->rp[i], v);
}

static double col_dot_dd(const pmat* M, int32_t j, const double* v) {
    return dd_dot_line(M->tv + M->cs[j], M->ri + M->cs[j], M->cs[j + 1] - M->cs[j], v);
}

static int cert_check(const pmat* M, const double* c, const double* lo, const double* up, const double* rl,
                      const double* ru, const double* dx, const double* dy, double eps, double* ratios) {
    dd pobj = {0, 0}, pabs = {0, 0}, pviol = {0, 0}, dobj = {0, 0}, dabs = {0, 0}, dviol = {0, 0};
    for (int32_t j = 0; j < M->n; ++j) {
        const double lam = -col_dot_dd(M, j, dy);
        const int lf = lo[j] > -INFINITY, uf = up[j] < INFINITY;
        double lp = lam;
        if (!lf) lp = smin(lp, 0.0);
        if (!uf) lp = smax(lp, 0.0);
        const double r = lam - lp;
        dd_add_prod(&dviol, r, r);
        if (lf && lp > 0.0) {
            dd_add_prod(&dobj, lo[j], lp);
            dd_add_prod(&dabs, fabs(lo[j]), lp);
        }
        if (uf && lp < 0.0) {
            dd_add_prod(&dobj, up[j], lp);
            dd_add_prod(&dabs, fabs(up[j]), -lp);
        }
        dd_add_prod(&pobj, c[j], dx[j]);
        dd_add_prod(&pabs, fabs(c[j]), fabs(dx[j]));
        double v = 0.0;
        if (lf && dx[j] < 0.0) v = -dx[j];
        if (uf && dx[j] > 0.0) v = dx[j];
        dd_add_prod(&pviol, v, v);
    }
    for (int32_t i = 0; i < M->m; ++i) {
        const double h = row_dot_dd(M, i, dx);
        const int lf = rl[i] > -INFINITY, uf = ru[i] < INFINITY;
        double v = 0.0;
        if (lf && h < 0.0) v = -h;
        if (uf && h > 0.0) v = h;
        dd_add_prod(&pviol, v, v);
        double sv = 0.0;
        if (!lf && dy[i] > 0.0) sv = dy[i];
        if (!uf && dy[i] < 0.0) sv = -dy[i];
        dd_add_prod(&dviol, sv, sv);
        if (lf && dy[i] > 0.0) {
            dd_add_prod(&dobj, rl[i], dy[i]);
            dd_add_prod(&dabs, fabs(rl[i]), dy[i]);
        }
        if (uf && dy[i] < 0.0) {
            dd_add_prod(&dobj, ru[i], dy[i]);
            dd_add_prod(&dabs, fabs(ru[i]), -dy[i]);
        }
    }
    const double D = ddv(dobj), Pn = -ddv(pobj);
    const double rd = (D > 0.0 && D > 1e-12 * ddv(dabs)) ? sqrt(ddv(dviol)) / D : INFINITY;
    const double rp = (Pn > 0.0 && Pn > 1e-12 * ddv(pabs)) ? sqrt(ddv(pviol)) / Pn : INFINITY;
    if (ratios) {
        ratios[0] = rd;
        ratios[1] = rp;
    }
    if (rd <= eps) return 21;
    if (rp <= eps) return 22;
    return 0;
}

/* CSR(A) beside the LP's CSC, both with their own (to be scaled) values. */
static void pmat_build(const or_lp* lp, pmat* Mt) {
    const int32_t m = lp->m, n = lp->n;
    const int64_t nnz = n > 0 ? lp->start[n] : 0;
    const size_t M1 = (size_t)(m > 0 ? m : 1), Z1 = (size_t)(nnz > 0 ? nnz : 1);
    Mt->m = m;
    Mt->n = n;
    Mt->rp = calloc(M1 + 1, sizeof(int32_t));
    Mt->ci = malloc(Z1 * sizeof(int32_t));
    Mt->av = malloc(Z1 * sizeof(double));
    Mt->cs = lp->start;
    Mt->ri = lp->index;
    Mt->tv = malloc(Z1 * sizeof(double));
    for (int64_t p = 0; p < nnz; ++p) {
        Mt->rp[lp->index[p] + 1]++;
        Mt->tv[p] = lp->value[p];
    }
    for (int32_t i = 0; i < m; ++i) Mt->rp[i + 1] += Mt->rp[i];
    int32_t* fill = malloc(M1 * sizeof(int32_t));
    for (int32_t i = 0; i < m; ++i) fill[i] = Mt->rp[i];
    for (int32_t j = 0; j < n; ++j)
        for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) {
            const int32_t q = fill[lp->index[p]]++;
            Mt->ci[q] = j;
            Mt->av[q] = lp->value[p];
        }
    free(fill);
}

static void pmat_free(pmat* Mt) {
    free(Mt->rp);
    free(Mt->ci);
    free(Mt->av);
    free(Mt->tv);
}

/* The pdlp mode's scaling, applied to Mt's values in place: `ruiz` passes of
 * exponent-domain Ruiz equilibration (factor 2^-floor((e+1)/2) of each row's
 * max |a|, e = floor(log2), then of each column's max |a| on the row-scaled
 * matrix), then `pc` Pock-Chambolle passes with alpha = 1 (the same with the
 * 1-norms sum |a| instead of the maxima). R (m) and C (n) receive the
 * accumulated factors, every one an exact power of two, so the scaled
 * values are a_ij * R_i * C_j exactly. */
static void pdlp_scale(pmat* Mt, const or_lp* lp, int ruiz, int pc, double* R, double* C) {
    const int32_t m = lp->m, n = lp->n;
    const int64_t nnz = n > 0 ? lp->start[n] : 0;
    const size_t M1 = (size_t)(m > 0 ? m : 1), N1 = (size_t)(n > 0 ? n : 1);
    for (int32_t i = 0; i < m; ++i) R[i] = 1.0;
    for (int32_t j = 0; j < n; ++j) C[j] = 1.0;
    double* f = malloc((M1 > N1 ? M1 : N1) * sizeof(double));
    for (int pass = 0; pass < ruiz + pc; ++pass) {
        const int sum = pass >= ruiz;
        for (int32_t i = 0; i < m; ++i) {
            double acc = 0.0;
            if (sum)
                acc = or_pdlp_row_sum(Mt->av + Mt->rp[i], NULL, (int64_t)Mt->rp[i + 1] - Mt->rp[i], NULL);
            else
                for (int32_t p = Mt->rp[i]; p < Mt->rp[i + 1]; ++p) acc = smax(acc, fabs(Mt->av[p]));
            f[i] = pow2_inv_sqrt(acc);
            R[i] *= f[i];
        }
        for (int32_t i = 0; i < m; ++i)
            for (int32_t p = Mt->rp[i]; p < Mt->rp[i + 1]; ++p) Mt->av[p] = Mt->av[p] * (f[i] * 1.0);
        for (int64_t p = 0; p < nnz; ++p) Mt->tv[p] = Mt->tv[p] * (1.0 * f[lp->index[p]]);
        for (int32_t j = 0; j < n; ++j) {
            double acc = 0.0;
            if (sum)
                acc = or_pdlp_row_sum(Mt->tv + lp->start[j], NULL, lp->start[j + 1] - lp->start[j], NULL);
            else
                for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) acc = smax(acc, fabs(Mt->tv[p]));
            f[j] = pow2_inv_sqrt(acc);
            C[j] *= f[j];
        }
        for (int32_t j = 0; j < n; ++j)
            for (int64_t p = lp->start[j]; p < lp->start[j + 1]; ++p) Mt->tv[p] = Mt->tv[p] * (f[j] * 1.0);
        for (int32_t i = 0; i < m; ++i)
            for (int32_t p = Mt->rp[i]; p < Mt->rp[i + 1]; ++p) Mt->av[p] = Mt->av[p] * (1.0 * f[Mt->ci[p]]);
    }
    free(f);
}

void or_scale_pdlp(const or_lp* lp, int32_t ruiz_passes, int32_t pc_passes, double* R, double* C) {
    pmat Mt;
    pmat_build(lp, &Mt);
    pdlp_scale(&Mt, lp, ruiz_passes, pc_passes, R, C);
    pmat_free(&Mt);
}

void or_pdlp(const or_lp* lp, const or_pdlp_params* prm, double* x_out, double* y_out, or_pdlp_info* info) {
    const int32_t m = lp->m, n = lp->n;
    const int64_t nnz = n > 0 ? lp->start[n] : 0;
    const size_t M1 = (size_t)(m > 0 ? m : 1), N1 = (size_t)(n > 0 ? n : 1);
    pmat Mt;
    pmat_build(lp, &Mt);
    /* scaling: 10 Ruiz passes, then one Pock-Chambolle pass */
    double* R = malloc(M1 * sizeof(double));
    double* C = malloc(N1 * sizeof(double));
    pdlp_scale(&Mt, lp, 10, 1, R, C);
    /* vectors in the scaled space */
    double *c = malloc(N1 * 8), *lo = malloc(N1 * 8), *up = malloc(N1 * 8), *ct = malloc(N1 * 8);
    double *rl = malloc(M1 * 8), *ru = malloc(M1 * 8), *rt = malloc(M1 * 8);
    dd bn = {0, 0}, cn = {0, 0};
    for (int32_t j = 0; j < n; ++j) {
        c[j] = lp->c[j] * C[j];
        lo[j] = fin(lp->l[j]) ? lp->l[j] / C[j] : -INFINITY;
        up[j] = fin(lp->u[j]) ? lp->u[j] / C[j] : INFINITY;
        const double us = prm->unscale_col ? prm->unscale_col[j] : 1.0;
        ct[j] = C[j] / us;
        const double cu = lp->c[j] / us;
        dd_add_prod(&cn, cu, cu);
    }
    for (int32_t i = 0; i < m; ++i) {
        rl[i] = fin(lp->rl[i]) ? lp->rl[i] * R[i] : -INFINITY;
        ru[i] = fin(lp->ru[i]) ? lp->ru[i] * R[i] : INFINITY;
        const double us = prm->unscale_row ? prm->unscale_row[i] : 1.0;
        rt[i] = R[i] * us;
        double b = 0.0;
        if (fin(lp->rl[i])) b = fabs(lp->rl[i]);
        if (fin(lp->ru[i])) b = smax(b, fabs(lp->ru[i]));
        b = b / us;
        dd_add_prod(&bn, b, b);
    }
    /* start point */
    double *x = malloc(N1 * 8), *xa = malloc(N1 * 8), *xp = malloc(N1 * 8), *xb = malloc(N1 * 8);
    double *y = malloc(M1 * 8), *ya = malloc(M1 * 8), *yp = malloc(M1 * 8);
    double *dxr = malloc(N1 * 8), *dyr = malloc(M1 * 8);
    for (int32_t j = 0; j < n; ++j) {
        double v = prm->x0 ? prm->x0[j] / C[j] : 0.0;
        v = (lo[j] < v) ? v : lo[j];
        v = (v < up[j]) ? v : up[j];
        x[j] = xa[j] = xp[j] = xb[j] = v;
    }
    for (int32_t i = 0; i < m; ++i) y[i] = ya[i] = yp[i] = prm->y0 ? prm->y0[i] / R[i] : 0.0;
    /* power iteration on A'^T A' */
    double lam = 0.0;
    if (n > 0 && m > 0 && nnz > 0) {
        double* u = malloc(N1 * 8);
        double* v = malloc(M1 * 8);
        for (int32_t j = 0; j < n; ++j) u[j] = 1.0 + (double)(((unsigned)j * 2654435761u) % 1000u) / 1000.0;
        dd s2 = {0, 0};
        for (int32_t j = 0; j < n; ++j) s2 = dd_sum(s2, ddp(u[j], u[j]));
        double nu = sqrt(ddv(s2));
        if (nu > 0)
            for (int32_t j = 0; j < n; ++j) u[j] = u[j] / nu;
        for (int it = 0; it < 200; ++it) {
            for (int32_t i = 0; i < m; ++i) v[i] = row_dot(&Mt, i, u);
            for (int32_t j = 0; j < n; ++j) u[j] = col_dot(&Mt, j, v);
            dd t = {0, 0};
            for (int32_t j = 0; j < n; ++j) t = dd_sum(t, ddp(u[j], u[j]));
            const double nw = sqrt(ddv(t));
            if (!(nw > 0.0) || !isfinite(nw)) break;
            for (int32_t j = 0; j < n; ++j) u[j] = u[j] / nw;
            const double prev = lam;
            lam = nw;
            if (it > 0 && fabs(lam - prev) <= 1e-7 * lam) break;
        }
        free(u);
        free(v);
    }
    const double smx = lam > 0.0 ? sqrt(lam) : 1.0;
    pstate s;
    memset(&s, 0, sizeof(s));
    s.eta = 0.998 / smx;
    {
        dd cs = {0, 0}, bs = {0, 0};
        for (int32_t j = 0; j < n; ++j) dd_add_prod(&cs, c[j], c[j]);
        for (int32_t i = 0; i < m; ++i) {
            double b = 0.0;
            if (isfinite(rl[i])) b = fabs(rl[i]);
            if (isfinite(ru[i])) b = smax(b, fabs(ru[i]));
            dd_add_prod(&bs, b, b);
        }
        const double cnorm = sqrt(cs.hi + cs.lo), bnorm = sqrt(bs.hi + bs.lo);
        s.omega = (cnorm > 1e-10 && bnorm > 1e-10) ? cnorm / bnorm : 1.0;
    }
    s.tau = s.eta / s.omega;
    s.sigma = s.eta * s.omega;
    s.tol = prm->tol;
    s.bnorm_o = sqrt(bn.hi + bn.lo);
    s.cnorm_o = sqrt(cn.hi + cn.lo);
    s.offset = lp->offset;
    s.max_iters = prm->max_iters > 0 ? prm->max_iters : 1;
    s.eval_freq = prm->eval_freq > 0 ? prm->eval_freq : 64;
    pdlp_eval(&Mt, &s, c, lo, up, rl, ru, rt, ct, xp, yp, xa, ya, 1);
    while (!s.done) {
        for (int kl = 0; kl < s.eval_freq; ++kl) {
            const int k = s.k_base + kl;
            const double w = (double)(k + 1) / (double)(k + 2);
            const double wc = 1.0 / (double)(k + 2);
            const int save = kl == s.eval_freq - 1;
            for (int32_t j = 0; j < n; ++j) {
                const double g = col_dot(&Mt, j, y);
                const double xj = x[j];
                double cand = xj - s.tau * (c[j] - g);
                cand = (lo[j] < cand) ? cand : lo[j];
                cand = (cand < up[j]) ? cand : up[j];
                const double b = 2.0 * cand - xj;
                xb[j] = b;
                x[j] = w * b + wc * xa[j];
                if (save) xp[j] = cand;
            }
            for (int32_t i = 0; i < m; ++i) {
                const double h = row_dot(&Mt, i, xb);
                const double yi = y[i];
                double r = h - yi / s.sigma;
                r = (rl[i] < r) ? r : rl[i];
                r = (r < ru[i]) ? r : ru[i];
                const double ypi = yi + s.sigma * (r - h);
                const double b = 2.0 * ypi - yi;
                y[i] = w * b + wc * ya[i];
                if (save) yp[i] = ypi;
            }
        }
        pdlp_eval(&Mt, &s, c, lo, up, rl, ru, rt, ct, xp, yp, xa, ya, 0);
        if (s.restart) {
            for (int32_t j = 0; j < n; ++j) dxr[j] = xp[j] - xa[j];
            for (int32_t i = 0; i < m; ++i) dyr[i] = yp[i] - ya[i];
            for (int32_t j = 0; j < n; ++j) x[j] = xa[j] = xp[j];
            for (int32_t i = 0; i < m; ++i) y[i] = ya[i] = yp[i];
            double ratios[2];
            const int cs = cert_check(&Mt, c, lo, up, rl, ru, dxr, dyr, 1e-8, ratios);
            if (getenv("OR_CERT_DEBUG"))
                fprintf(stderr, "restart %d at %lld: dual-ray ratio %.3e primal-ray ratio %.3e\n", s.restarts,
                        (long long)s.total, ratios[0], ratios[1]);
            if (cs) {
                s.done = 1;
                s.status = cs;
            }
        }
    }
    for (int32_t j = 0; j < n; ++j) x_out[j] = xp[j] * C[j];
    for (int32_t i = 0; i < m; ++i) y_out[i] = yp[i] * R[i];
    if (info) {
        info->status = s.status;
        info->iterations = s.total;
        info->restarts = s.restarts;
        info->omega = s.omega;
        info->eta = s.eta;
        info->rel_p = s.rel_p;
        info->rel_d = s.rel_d;
        info->rel_g = s.rel_g;
        info->pobj = s.pobj;
        info->dobj = s.dobj;
    }
    pmat_free(&Mt); free(R); free(C);
    free(c); free(lo); free(up); free(ct); free(rl); free(ru); free(rt);
    free(x); free(xa); free(xp); free(xb); free(y); free(ya); free(yp); free(dxr); free(dyr);
}
