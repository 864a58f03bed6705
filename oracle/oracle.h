/* ORACLE - test infrastructure only.
 *
 * A plain-C (C99) restatement of the reference's CPU PDHG path, used as the
 * checker for the GPU engine. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only to check or time against
 * it; the product never calls into this directory.
 *
 * Compiled with gcc -O3 -ffp-contract=off (no fused multiply-add), matching
 * the rounding of the reference's plain x86-64 build (Makefile:2).
 *
 * Function -> reference lines it restates -> what pins it
 *   or_spmv          src/core/csc_matrix.hpp:99-108   KAT unit_tests.cpp:56-60,
 *                                                     dense brute force, oracle/_ref
 *   or_spmv_t        src/core/csc_matrix.hpp:111-120  dense brute force, oracle/_ref
 *   or_scale_geometric src/model/scaling.hpp:14-154   pow2 / equilibrium invariants,
 *                                                     oracle/_ref factors
 *   or_pdhg_reference src/gpu/pdhg_solver.hpp:25-107  closed-form step sizes and first
 *                                                     step of unit_tests.cpp:316-331,
 *                                                     oracle/_ref iterates at K >= 5000
 *   or_pdhg_steps    the loop of :59-101 in chunks    chunked run = uninterrupted run
 *   or_dd_verify     src/io/sol_io.hpp:25-106         KAT unit_tests.cpp:29-42,
 *                                                     README.md:84 (2.55e-14)
 *   or_kkt           build-defined (PDLP relative     zero at the A.5 optimum; closed-form
 *                    KKT, caller units)               nonzero terms at perturbed points of
 *                                                     the A.5 and unit_tests.cpp:316-331 LPs
 *   or_scale_pdlp    build-defined (Ruiz + Pock-      exact powers of two; inf-norm
 *                    Chambolle)                       equilibrium after the Ruiz passes;
 *                                                     hand-computed factors; the final
 *                                                     pass = one PC pass on the Ruiz result
 *   or_pdlp_row_sum  build-defined product order      closed forms on rows whose tiny
 *                    (rows > 4096 chunked 64 / 32;    terms survive only per chunk and
 *                    also the Pock-Chambolle measure  per group (test_oracle_pdlp.py);
 *                    and, in dd, the certificate)     HiGHS verdicts on rays through
 *                                                     long rows and columns
 *   or_pdlp          build-defined restarted PDHG     KKT = 0 at known optima,
 *                                                     scaling invariance, HiGHS optimum;
 *                    + ray certificates at restarts   status 21 / 22 exactly on LPs
 *                                                     scipy's HiGHS finds infeasible /
 *                                                     unbounded, 0 on feasible ones
 */
#ifndef NIR_ORACLE_H
#define NIR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LP in the reference's CSC layout: min c^T x + off, rl <= A x <= ru, l <= x <= u. */
typedef struct {
    int32_t m, n;
    const int64_t* start;
    const int32_t* index;
    const double* value;
    const double* c;
    const double* l;
    const double* u;
    const double* rl;
    const double* ru;
    double offset;
} or_lp;

void or_spmv(const or_lp* lp, const double* x, double* y);   /* y = A x  (m) */
void or_spmv_t(const or_lp* lp, const double* y, double* g); /* g = A^T y (n) */

/* Geometric-mean scaling factors; returns 1 when scaling engages. */
int or_scale_geometric(const or_lp* lp, double* R, double* C);

/* Reference PDHG. Runs at most K iterations; with check != 0 it applies the
 * reference stop rule. x (n) and y (m) receive the final iterates; the
 * remaining outputs may be NULL. */
void or_pdhg_reference(const or_lp* lp, int64_t K, int check, const double* x0, const double* y0, double* x,
                       double* y, double* tau_out, int64_t* iters_out, double* viol_out);

/* K iterations of the reference loop (no stop rule) from (x, y) in place,
 * with the given step sizes: continuing a run in chunks gives the iterates of
 * one uninterrupted run (the loop keeps no state beyond x and y). */
void or_pdhg_steps(const or_lp* lp, int64_t K, double tau, double sigma, double* x, double* y);

/* Double-double checker: returns 1 on pass. */
int or_dd_verify(const or_lp* lp, const double* x, double claimed_obj, double tol, double* max_row_viol,
                 double* max_bound_viol, double* computed_obj);

/* Relative KKT of (x, y) on the LP (see or_pdlp). out[0..2] = relative primal,
 * dual, gap; out[3], out[4] = primal and dual objective. */
void or_kkt(const or_lp* lp, const double* x, const double* y, double* out);

typedef struct {
    int64_t max_iters;
    double tol;
    int32_t eval_freq;
    const double* unscale_row; /* may be NULL */
    const double* unscale_col; /* may be NULL */
    const double* x0;          /* may be NULL */
    const double* y0;          /* may be NULL */
} or_pdlp_params;

typedef struct {
    int32_t status; /* 0 optimal, 6 iteration limit, 11 numerical trouble,
                       21 a dual ray certifies infeasibility,
                       22 a primal ray certifies dual infeasibility */
    int32_t restarts;
    int64_t iterations;
    double omega, eta, rel_p, rel_d, rel_g, pobj, dobj;
} or_pdlp_info;

/* The pdlp mode's scaling factors: `ruiz_passes` exponent-domain Ruiz passes
 * (rows by max |a|, then columns of the row-scaled matrix), then `pc_passes`
 * Pock-Chambolle passes (alpha = 1: 1-norms). The engine uses 10 + 1.
 * R (m) and C (n) are exact powers of two; scaled A' = diag(R) A diag(C). */
void or_scale_pdlp(const or_lp* lp, int32_t ruiz_passes, int32_t pc_passes, double* R, double* C);

/* pdlp mode's sum of one row's (or column's) terms val[p] * v[idx[p]]
 * (|val[p]| when v is NULL: the Pock-Chambolle row measure): sequential from
 * 0.0 for len <= 4096; above that, chunks of 64 positions summed from 0.0,
 * the chunk sums from 0.0 in groups of 32, the group sums from 0.0 into the
 * result. The certificate's double-double products of lines above 4096
 * nonzeros follow the same chunks and groups in double-double. */
double or_pdlp_row_sum(const double* val, const int32_t* idx, int64_t len, const double* v);

/* The engine's pdlp mode in FP64 (see pdhg_oracle.c). x (n) and y (m) are
 * returned in the LP's own units. */
void or_pdlp(const or_lp* lp, const or_pdlp_params* p, double* x, double* y, or_pdlp_info* info);

#ifdef __cplusplus
}
#endif
#endif
