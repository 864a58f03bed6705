// Arithmetic of the restarted PDHG ("pdlp" mode) shared by the single-LP
// kernels (pdlp.hip) and the batched persistent kernel (batch.hip): the same
// inline functions produce every iterate, evaluation term and decision in
// both, so a batched instance follows the single-LP trajectory exactly.
#pragma once
#include <hip/hip_runtime.h>

#include "spmv.hpp"

namespace nir {
namespace gpu {

// ------------------------------------------------------------------ double-double

struct dd {
    double hi, lo;
};

__host__ __device__ __forceinline__ dd dd_zero() { return dd{0.0, 0.0}; }

// Sum of two dd values (two-sum on both parts, then renormalise).
__host__ __device__ __forceinline__ dd dd_add(dd a, dd b) {
    double s1 = a.hi + b.hi, v = s1 - a.hi;
    double s2 = (a.hi - (s1 - v)) + (b.hi - v);
    double t1 = a.lo + b.lo, w = t1 - a.lo;
    double t2 = (a.lo - (t1 - w)) + (b.lo - w);
    s2 += t1;
    double u = s1 + s2;
    s2 = s2 - (u - s1);
    s1 = u;
    s2 += t2;
    dd r;
    r.hi = s1 + s2;
    r.lo = s2 - (r.hi - s1);
    return r;
}

// Exact product a*b as a dd (FMA two-product).
__host__ __device__ __forceinline__ dd dd_prod(double a, double b) {
    const double p = a * b;
    return dd{p, fma(a, b, -p)};
}

__host__ __device__ __forceinline__ double dd_val(dd a) { return a.hi + a.lo; }

// sum_p val[p] v[col[p]] over row i of a segmented CSR (row pointers
// nseg * (nrows + 1)), in ascending column order (segment by segment), in
// double-double.
__device__ __forceinline__ dd row_dot_dd(int nrows, int nseg, const int* __restrict__ rp,
                                         const int* __restrict__ col, const double* __restrict__ val,
                                         const double* __restrict__ v, int i) {
    dd acc = dd_zero();
    for (int s = 0; s < nseg; ++s) {
        const int* r = rp + static_cast<size_t>(s) * (nrows + 1);
        for (int p = r[i]; p < r[i + 1]; ++p) acc = dd_add(acc, dd_prod(val[p], v[col[p]]));
    }
    return acc;
}

// 2^-floor((e+1)/2) for v in [2^e, 2^(e+1)); 1 for v == 0 or non-finite.
__host__ __device__ __forceinline__ double pow2_inv_sqrt(double v) {
    if (!(v > 0.0) || !isfinite(v)) return 1.0;
    int e2;
    (void)frexp(v, &e2);   // v = f 2^e2, f in [0.5, 1)
    const int e = e2 - 1;
    const int k = (e + 1) >= 0 ? (e + 1) / 2 : -((-(e + 1) + 1) / 2);   // floor((e+1)/2)
    return ldexp(1.0, -k);
}

// ------------------------------------------------------------------ state

struct PdlpState {
    double omega, tau, sigma, eta;
    double kkt_anchor, kkt_prev;
    double rel_p, rel_d, rel_g, pobj, dobj;
    double best_rel;
    double tol;
    double bnorm_o, cnorm_o, offset;
    long long total;       // iterations completed
    long long max_iters;
    int k_base;            // iterations since restart at the start of the block
    int done;              // 1 = stop iterating
    int status;            // 0 optimal, 6 iteration limit, 11 numerical trouble, 2 stalled (mixed phase)
    int restart;           // set by the decision for the restart copy
    int restarts;
    int evals;
    int since_best;        // evaluations since best_rel last improved by 5%
    int watch_stall;       // 1 in the FP32 phase
    int eval_freq;
};

// ------------------------------------------------------------------ iteration

// Halpern weights of inner iteration k (iterations since the last restart):
// z+ = w (2 T(z) - z) + wc z_anchor.
template <typename T>
struct Halpern {
    T w, wc;
};
template <typename T>
__device__ __forceinline__ Halpern<T> halpern(int k) {
    return Halpern<T>{static_cast<T>(static_cast<double>(k + 1) / static_cast<double>(k + 2)),
                      static_cast<T>(1.0 / static_cast<double>(k + 2))};
}

// Projected primal step at column j from g = (A^T y)_j: T(z)_x = cand and the
// reflected point xb = 2 cand - x (the extrapolated point of the dual step).
template <typename T>
struct PrimalOut {
    T cand, xb, next;
};
template <typename T>
__device__ __forceinline__ PrimalOut<T> primal_point(T xj, T g, T c, T l, T u, T xa, T tau, Halpern<T> h) {
    T cand = xj - tau * (c - g);
    cand = (l < cand) ? cand : l;
    cand = (cand < u) ? cand : u;
    const T xb = T(2) * cand - xj;
    return PrimalOut<T>{cand, xb, h.w * xb + h.wc * xa};
}

// Projected dual step at row i from hx = (A xb)_i.
template <typename T>
struct DualOut {
    T yp, next;
};
template <typename T>
__device__ __forceinline__ DualOut<T> dual_point(T yi, T hx, T l, T u, T ya, T sigma, Halpern<T> h) {
    T r = hx - yi / sigma;
    r = (l < r) ? r : l;
    r = (r < u) ? r : u;
    const T yp = yi + sigma * (r - hx);
    const T yb = T(2) * yp - yi;
    return DualOut<T>{yp, h.w * yb + h.wc * ya};
}

// Re-centred FP32 iteration (mixed precision). The iterate is an FP64 anchor
// (X, Y) plus FP32 deltas (dx, dy); the anchor is also the Halpern anchor, so
// its delta is 0 and the Halpern step is next = w * (2 T(dz) - dz). At every
// restart the anchor moves to T(z) in FP64 and the FP32 constants are
// re-based on it: rc = c - A^T Y, [dl, du] = [l - X, u - X],
// [drl, dru] = [rl - A X, ru - A X], sa = -Y / sigma. The FP32 quantities then
// scale with the distance to the anchor, not with the solution, so FP32
// rounding shrinks as the iterates converge.
//
// Primal: primal_point with c = rc, bounds [dl, du] and a zero anchor.
// Dual, from h = (A dxb)_i:
//     r = clamp(h - dy/sigma + sa, drl, dru);  dyp = dy + sigma (r - h).
template <typename T>
__device__ __forceinline__ DualOut<T> dual_point_delta(T dy, T h, T drl, T dru, T sa, T sigma, Halpern<T> hw) {
    T r = h - dy / sigma + sa;
    r = (drl < r) ? r : drl;
    r = (r < dru) ? r : dru;
    const T yp = dy + sigma * (r - h);
    const T yb = T(2) * yp - dy;
    return DualOut<T>{yp, hw.w * yb};
}

// ------------------------------------------------------------------ evaluation

// Column terms at T(z) with g = (A^T yp)_j: 0 rd^2, 1 (rd/C)^2, 2 c x,
// 3 dual-objective column term, 4 (x - x_anchor)^2.
__device__ __forceinline__ void col_terms(double g, double c, double lo, double up, double cscale, double xj,
                                          double xaj, dd (&t)[5]) {
    const double lam = c - g;
    const bool lf = lo > -INFINITY, uf = up < INFINITY;
    double lp = lam;
    if (!lf) lp = amin(lp, 0.0);
    if (!uf) lp = amax(lp, 0.0);
    const double rd = lam - lp;
    t[0] = dd_prod(rd, rd);
    const double rdo = rd / cscale;
    t[1] = dd_prod(rdo, rdo);
    t[2] = dd_prod(c, xj);
    dd dobj = dd_zero();
    if (lf && lp > 0.0) dobj = dd_prod(lo, lp);
    if (uf && lp < 0.0) dobj = dd_prod(up, lp);
    t[3] = dobj;
    const double dx = xj - xaj;
    t[4] = dd_prod(dx, dx);
}

// Row terms at T(z) with h = (A xp)_i: 0 rp^2, 1 (rp/R)^2, 2 yviol^2,
// 3 (R yviol)^2, 4 dual-objective row term, 5 (y - y_anchor)^2.
__device__ __forceinline__ void row_terms(double h, double rl, double ru, double rscale, double yi, double yai,
                                          dd (&t)[6]) {
    const bool lf = rl > -INFINITY, uf = ru < INFINITY;
    double proj = amax(rl, h);
    proj = amin(ru, proj);
    const double rp = h - proj;
    t[0] = dd_prod(rp, rp);
    const double rpo = rp / rscale;
    t[1] = dd_prod(rpo, rpo);
    double ys = yi;
    if (!lf) ys = amin(ys, 0.0);
    if (!uf) ys = amax(ys, 0.0);
    const double yv = yi - ys;
    t[2] = dd_prod(yv, yv);
    const double yvo = yv * rscale;
    t[3] = dd_prod(yvo, yvo);
    dd dobj = dd_zero();
    if (lf && ys > 0.0) dobj = dd_prod(rl, ys);
    if (uf && ys < 0.0) dobj = dd_prod(ru, ys);
    t[4] = dobj;
    const double dy = yi - yai;
    t[5] = dd_prod(dy, dy);
}

// ------------------------------------------------------------------ decision

// KKT evaluation and the restart / primal-weight / termination decision from
// the 11 reduced terms (5 column terms, then 6 row terms). mode 0: regular
// evaluation after a block of eval_freq iterations; mode 1: evaluation of the
// start point (sets the anchor KKT, may terminate).
__device__ __forceinline__ void pdlp_decide(PdlpState& s, const dd* tot, int mode) {
    const double rd2 = dd_val(dd_add(tot[0], tot[7]));
    const double rd2o = dd_val(dd_add(tot[1], tot[8]));
    const double rp2 = dd_val(tot[5]);
    const double rp2o = dd_val(tot[6]);
    const double P = dd_val(dd_add(tot[2], dd{s.offset, 0.0}));
    const double D = dd_val(dd_add(dd_add(tot[3], tot[9]), dd{s.offset, 0.0}));
    const double dx2 = dd_val(tot[4]);
    const double dy2 = dd_val(tot[10]);
    const double gap = fabs(P - D);
    s.rel_p = sqrt(rp2o) / (1.0 + s.bnorm_o);
    s.rel_d = sqrt(rd2o) / (1.0 + s.cnorm_o);
    s.rel_g = gap / (1.0 + fabs(P) + fabs(D));
    s.pobj = P;
    s.dobj = D;
    s.evals += 1;
    const double worst = amax(s.rel_p, amax(s.rel_d, s.rel_g));
    if (!isfinite(rd2) || !isfinite(rp2) || !isfinite(P) || !isfinite(D) || !isfinite(dx2) || !isfinite(dy2)) {
        s.done = 1;
        s.status = 11;
        s.restart = 0;
        return;
    }
    const double kkt = sqrt(s.omega * rp2 + rd2 / s.omega + gap * gap);
    if (mode == 1) {
        s.kkt_anchor = kkt;
        s.kkt_prev = kkt;
        s.best_rel = worst;
        s.since_best = 0;
        s.restart = 0;
        if (worst <= s.tol) {
            s.done = 1;
            s.status = 0;
        }
        return;
    }
    s.total += s.eval_freq;
    if (worst <= s.tol) {
        s.done = 1;
        s.status = 0;
        s.restart = 0;
        return;
    }
    if (s.watch_stall) {
        // FP32 phase: hand over to FP64 once the best relative KKT has not
        // improved by 5% for max(100, evals/3) evaluations.
        if (worst <= 0.95 * s.best_rel) {
            s.best_rel = worst;
            s.since_best = 0;
        } else if (++s.since_best >= (s.evals / 3 > 100 ? s.evals / 3 : 100)) {
            s.done = 1;
            s.status = 2;
            s.restart = 0;
            return;
        }
    }
    if (s.total >= s.max_iters) {
        s.done = 1;
        s.status = 6;
        s.restart = 0;
        return;
    }
    const int k_new = s.k_base + s.eval_freq;
    const bool restart = (kkt <= 0.2 * s.kkt_anchor) || (kkt <= 0.8 * s.kkt_anchor && kkt > s.kkt_prev) ||
                         (static_cast<double>(k_new) >= 0.36 * static_cast<double>(s.total));
    s.kkt_prev = kkt;
    if (restart) {
        if (dx2 > 1e-20 && dy2 > 1e-20) {
            const double ratio = sqrt(dy2) / sqrt(dx2);
            s.omega = sqrt(s.omega * ratio);
            s.tau = s.eta / s.omega;
            s.sigma = s.eta * s.omega;
        }
        s.kkt_anchor = sqrt(s.omega * rp2 + rd2 / s.omega + gap * gap);
        s.kkt_prev = s.kkt_anchor;
        s.k_base = 0;
        s.restarts += 1;
        s.restart = 1;
    } else {
        s.k_base = k_new;
        s.restart = 0;
    }
}

// ------------------------------------------------------------------ infeasibility certificates

// On an infeasible or unbounded LP the restarted iterates drift along a ray,
// so at every restart the step between the old and the new anchor, d = (dx,
// dy), is tested as a certificate in the engine's scaled space:
//   dual ray dy (primal infeasible): lam = -A^T dy; violations lam - proj(lam)
//     for infinite column bounds and dy_i > 0 / < 0 for infinite rl_i / ru_i;
//     objective D = sum rl_i dy_i [dy_i > 0] + ru_i dy_i [dy_i < 0]
//                 + sum l_j lp_j [lp_j > 0] + u_j lp_j [lp_j < 0] > 0;
//   primal ray dx (dual infeasible): the parts of A dx and dx outside the
//     recession cones of the rows and bounds; objective -c^T dx > 0.
// A dx and A^T dy are double-double dot products over the FP64 matrices and
// every sum is double-double. A ray passes when its objective exceeds 1e-12
// of the sum of its terms' magnitudes and ||violations||_2 <= kCertEps times
// the objective.
constexpr double kCertEps = 1e-8;
constexpr int kStPrimalInfeasible = 21;   // PdlpState::status: a dual ray passed
constexpr int kStDualInfeasible = 22;     // PdlpState::status: a primal ray passed

// Column j with lam = -(A^T dy)_j; slots: 0 dual violation^2, 1 D, 2 |D
// terms|, 3 c^T dx, 4 |c| |dx|, 5 primal violation^2.
__device__ __forceinline__ void cert_col_terms(double lam, double c, double lo, double up, double dx, dd (&acc)[6]) {
    const bool lf = lo > -INFINITY, uf = up < INFINITY;
    double lp = lam;
    if (!lf) lp = amin(lp, 0.0);
    if (!uf) lp = amax(lp, 0.0);
    const double r = lam - lp;
    acc[0] = dd_add(acc[0], dd_prod(r, r));
    if (lf && lp > 0.0) {
        acc[1] = dd_add(acc[1], dd_prod(lo, lp));
        acc[2] = dd_add(acc[2], dd_prod(fabs(lo), lp));
    }
    if (uf && lp < 0.0) {
        acc[1] = dd_add(acc[1], dd_prod(up, lp));
        acc[2] = dd_add(acc[2], dd_prod(fabs(up), -lp));
    }
    acc[3] = dd_add(acc[3], dd_prod(c, dx));
    acc[4] = dd_add(acc[4], dd_prod(fabs(c), fabs(dx)));
    double v = 0.0;
    if (lf && dx < 0.0) v = -dx;
    if (uf && dx > 0.0) v = dx;
    acc[5] = dd_add(acc[5], dd_prod(v, v));
}

// Row i with h = (A dx)_i, same slots (3 and 4 untouched).
__device__ __forceinline__ void cert_row_terms(double h, double rl, double ru, double dy, dd (&acc)[6]) {
    const bool lf = rl > -INFINITY, uf = ru < INFINITY;
    double v = 0.0;
    if (lf && h < 0.0) v = -h;
    if (uf && h > 0.0) v = h;
    acc[5] = dd_add(acc[5], dd_prod(v, v));
    double sv = 0.0;
    if (!lf && dy > 0.0) sv = dy;
    if (!uf && dy < 0.0) sv = -dy;
    acc[0] = dd_add(acc[0], dd_prod(sv, sv));
    if (lf && dy > 0.0) {
        acc[1] = dd_add(acc[1], dd_prod(rl, dy));
        acc[2] = dd_add(acc[2], dd_prod(fabs(rl), dy));
    }
    if (uf && dy < 0.0) {
        acc[1] = dd_add(acc[1], dd_prod(ru, dy));
        acc[2] = dd_add(acc[2], dd_prod(fabs(ru), -dy));
    }
}

// Decision from the six totals: kStPrimalInfeasible, kStDualInfeasible or 0.
__device__ __forceinline__ int cert_decide(const dd* tot) {
    const double D = dd_val(tot[1]), Pn = -dd_val(tot[3]);
    if (D > 0.0 && D > 1e-12 * dd_val(tot[2]) && sqrt(dd_val(tot[0])) <= kCertEps * D) return kStPrimalInfeasible;
    if (Pn > 0.0 && Pn > 1e-12 * dd_val(tot[4]) && sqrt(dd_val(tot[5])) <= kCertEps * Pn) return kStDualInfeasible;
    return 0;
}

// Status of a finished state in the reference's codes (src/core/status.hpp:7-24):
// a dual ray certifies infeasibility; a primal ray certifies dual
// infeasibility - unbounded when the last evaluated point met the primal
// tolerance, infeasible-or-unbounded otherwise.
__host__ __device__ inline int sov_status(const PdlpState& h) {
    if (h.status == kStPrimalInfeasible) return 1;
    if (h.status == kStDualInfeasible) return h.rel_p <= h.tol ? 2 : 3;
    return h.status;
}

}  // namespace gpu
}  // namespace nir
