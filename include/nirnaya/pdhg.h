/* PDHG engine C ABI: one MI355X, one HIP stream per context, device-resident
 * LP sessions.
 *
 * This ABI replaces the reference's internal operator
 *     sov::PdhgSolver::solve(const Model&, const Options&, const Deadline&)
 *         -> PdhgResult                         (src/gpu/pdhg_solver.hpp:14-25)
 * called from the facade at include/sov/solver.hpp:237. The LP arrives in the
 * reference's own storage (CSC with int64 column starts, int32 row indices,
 * double values; src/core/csc_matrix.hpp:18-28) plus the bound and cost
 * vectors of src/model/model.hpp:21-36, already presolved and scaled by the
 * caller. Conventions: int return codes (0 = success, -1 = failure with the
 * context's error string set), no exceptions across the boundary, caller-owned
 * host buffers, device buffers owned by the context / session.
 */
#ifndef NIRNAYA_PDHG_H
#define NIRNAYA_PDHG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NIR_API __attribute__((visibility("default")))

typedef struct nir_ctx nir_ctx;
typedef struct nir_session nir_session;

/* min c^T x + obj_offset  s.t.  rowlo <= A x <= rowup,  collo <= x <= colup.
 * Bounds with |b| >= 1e20 (or non-finite) are absent. */
typedef struct {
    int32_t nrow;
    int32_t ncol;
    const int64_t* col_start; /* ncol + 1 */
    const int32_t* row_index; /* nnz, strictly ascending within a column */
    const double* value;      /* nnz */
    const double* cost;       /* ncol */
    const double* collo;      /* ncol */
    const double* colup;      /* ncol */
    const double* rowlo;      /* nrow */
    const double* rowup;      /* nrow */
    double obj_offset;
} nir_lp;

enum { NIR_MODE_REFERENCE = 0, NIR_MODE_PDLP = 1 };
enum { NIR_PREC_FP64 = 0, NIR_PREC_MIXED = 1 };

typedef struct {
    int32_t mode;              /* NIR_MODE_* */
    int32_t precision;         /* NIR_PREC_* (reference mode is always FP64) */
    int64_t max_iters;         /* iteration cap (reference mode: the loop bound K) */
    double tol;                /* pdlp: relative KKT tolerance */
    double polish_tol;         /* mixed: FP64 polish target, 0 = no polish */
    int32_t eval_freq;         /* pdlp: iterations between KKT evaluations */
    int32_t check_termination; /* reference mode: 1 = stop on the reference rule
                                  (max primal violation <= 1e-6 at k>20, k%32==0),
                                  0 = run exactly max_iters iterations */
    int32_t use_graph;         /* capture iteration blocks in hipGraphs */
    int32_t verbose;
    int32_t tile_nnz;          /* SpMV tile: nonzeros per workgroup (1024, 2048 or 4096;
                                  -1 = autotune tile, rows and segments per matrix; 0 = autotune
                                  when nnz >= 2^21 and seg_width is 0, else 2048) */
    int32_t tile_rows;         /* SpMV tile: max rows per workgroup (0 = default 256) */
    int32_t seg_width;         /* column-segment width of the A xbar products (0 = one segment;
                                  pdlp mode keeps one segment for a matrix with a row of more
                                  than 4,096 nonzeros, whose products it sums in chunks) */
    int32_t seg_width_t;       /* the same for the A^T y products (-1 = seg_width) */
    const double* x0;          /* optional warm start (ncol), NULL = cold */
    const double* y0;          /* optional warm start (nrow), NULL = cold */
    const double* unscale_row; /* optional: report KKT in units y*row, (Ax)/row */
    const double* unscale_col; /* optional: x*col, (c - A^T y)/col */
    int (*should_stop)(void* user); /* polled every 256 iterations; nonzero = stop */
    void* stop_user;
    double primal_weight;      /* pdlp: initial primal weight in the engine's scaled space
                                  (0 = ||c|| / ||b|| of the scaled LP); a warm start can carry
                                  the weight its family's solves ended with */
    int32_t seg_fused;         /* reference mode, segmented products: walk all segments in one
                                  launch with register running sums (1) or one launch per
                                  segment (0); the autotuner chooses when tile_nnz = -1 */
} nir_params;

typedef struct {
    int32_t status;         /* sov Status code: 0 optimal, 4 time limit, 6 iteration limit,
                               11 numerical trouble */
    int32_t restarts;
    int64_t iterations;     /* iterations completed (reference semantics) */
    int64_t polish_iterations;
    double objective;       /* c^T x + obj_offset of the returned x (double-double) */
    double tau, sigma;      /* step sizes in effect at the end */
    double primal_weight;
    double max_primal_viol; /* reference mode: last max row violation of A xbar */
    double rel_primal, rel_dual, rel_gap; /* last relative KKT terms (pdlp) */
    double setup_seconds;   /* upload + preprocessing */
    double solve_seconds;   /* iteration loop */
} nir_info;

/* Timing of `iterations` iterations, with HIP events on the session's stream.
 * Reference mode runs two sweeps of that many iterations: the production path
 * (32-iteration graph replays) between one event pair gives total_ms; then the
 * same kernels launched one by one with an event pair around each product
 * (all its column-segment passes) give spmv_ms / spmtv_ms. Pdlp mode runs one
 * sweep with an event pair around each product. */
typedef struct {
    int64_t iterations;
    double total_ms;            /* events bracketing the production sweep */
    double spmv_ms;             /* summed durations: A xbar + dual step */
    double spmtv_ms;            /* summed durations: A^T y + primal step */
    double spmv_bytes;          /* algorithmic bytes of one A xbar launch */
    double spmtv_bytes;         /* algorithmic bytes of one A^T y launch */
    double iter_bytes;          /* algorithmic bytes of one whole iteration */
    int32_t spmv_passes;        /* kernel dispatches per A xbar (segment passes, 1 when fused) */
    int32_t spmtv_passes;       /* kernel dispatches per A^T y */
} nir_timing;

NIR_API void nir_params_default(nir_params* p);
/* Layout check for bindings: {sizeof, offsetof(last field)} of nir_lp,
 * nir_params, nir_info, nir_timing, nir_batch_opts, nir_batch_stats,
 * nir_verify_report (14 values); returns the count, writes them when cap
 * allows. */
NIR_API int nir_abi_layout(int64_t* out, int32_t cap);

NIR_API int nir_ctx_create(int device, nir_ctx** out);
NIR_API void nir_ctx_destroy(nir_ctx* ctx);
NIR_API const char* nir_ctx_error(const nir_ctx* ctx);
NIR_API const char* nir_last_error(void); /* errors before a context exists */

NIR_API int nir_session_create(nir_ctx* ctx, const nir_lp* lp, const nir_params* p, nir_session** out);
/* Runs at most `iters` more iterations (stops earlier on termination). */
NIR_API int nir_session_run(nir_session* s, int64_t iters);
/* Runs to termination, the iteration cap or should_stop. */
NIR_API int nir_session_solve(nir_session* s);
/* Copies the current iterate (x: ncol, y: nrow; either may be NULL). */
NIR_API int nir_session_result(nir_session* s, double* x, double* y, nir_info* info);
NIR_API int nir_session_time(nir_session* s, int64_t iters, nir_timing* out);
/* Tile shapes in use: {A tile, A rows, A segments, A fused, A^T tile, A^T rows,
 * A^T segments, A^T fused} (fused = all segments walked in one launch). An A
 * tile of 0 marks a reference-mode session on the small-LP path: every
 * iteration in one workgroup with x, xbar and y in LDS (chosen for LPs of at
 * most 16,384 nonzeros, 2n + m <= 19,200, rows of A and A^T of at most 512
 * nonzeros, when no tile or segment width is given). */
NIR_API int nir_session_shape(nir_session* s, int32_t* out8);
/* Rows split over workgroups (pdlp mode, rows of more than 4,096 nonzeros in a
 * one-segment layout): out4 = {split rows of A, their pieces, split rows of
 * A^T (columns of A), their pieces}; all 0 in reference mode. */
NIR_API int nir_session_split(nir_session* s, int32_t* out4);
/* Relative KKT error of a reference-mode session's current iterate (x, y) on
 * its LP, evaluated on the device with double-double products and sums:
 * out5 = {relative primal, relative dual, relative gap, primal objective, dual
 * objective}, the definitions the pdlp mode reports in nir_info. The
 * reference loop itself stops only on its row-violation rule
 * (src/gpu/pdhg_solver.hpp:97-100); this lets a caller time the plain PDHG to
 * a KKT tolerance by running it in chunks (nir_session_run) between
 * evaluations. No counterpart in the reference. pdlp sessions: error (their
 * terms are in nir_info). */
NIR_API int nir_session_kkt(nir_session* s, double* out5);
/* Access floor of a reference-mode session's product (which = 0: A xbar,
 * 1: A^T y): microseconds per product of a kernel that makes only the
 * product's memory accesses over the session's own layout - each stored
 * nonzero's value and column index streamed, its operand entry gathered once,
 * one launch per column segment - averaged over `reps` products. The ratio
 * to the product's own duration (nir_session_time) shows what the row
 * structure, the segment passes and the fused epilogue add. No counterpart
 * in the reference. */
NIR_API int nir_session_gather_floor(nir_session* s, int32_t which, int32_t reps, double* us);
/* pdlp sessions: the engine's scaling factors, A' = diag(R) A diag(C) (R: nrow,
 * C: ncol; either may be NULL) - 10 exponent-domain Ruiz passes and one
 * Pock-Chambolle pass, every factor a power of two. This is the engine's own
 * equilibration, applied after (and independent of) the caller's
 * src/model/scaling.hpp:56-154 geometric scaling; no counterpart in the
 * reference. Reference sessions: error. */
NIR_API int nir_session_scaling(nir_session* s, double* R, double* C);
NIR_API void nir_session_destroy(nir_session* s);

/* Copies `bytes` (read) into a second buffer (write) with a plain grid-stride
 * kernel, k_copy_calibrate: the known-size reference for HBM byte counters. */
NIR_API int nir_calibrate_copy(nir_ctx* ctx, int64_t bytes, double* ms);
/* Random-gather rate: microseconds per launch of a grid-stride kernel
 * (`blocks` x 256 lanes) that gathers an FP64 operand of `len` entries at n
 * uniform random int32 indices (streamed); with_values 1 also multiplies
 * each by a streamed FP64 value, 2 streams the indices and values without
 * gathering; average of 10 launches after 2 warm ones. Measures the access
 * pattern of a sparse product over an unstructured matrix. */
NIR_API int nir_calibrate_gather(nir_ctx* ctx, int64_t n, int32_t len, int32_t with_values, int32_t blocks,
                                 double* us);

/* One-shot: create, solve, copy out, destroy. In pdlp mode an LP of at most
 * 2,500 nonzeros and 8,192 rows and columns, with no explicit layout and no
 * unscaling vectors, runs as a batch of one (nir_batch_create_warm, 256
 * iterations per launch): the same bits as the session path, one persistent
 * workgroup instead of two launches per iteration. */
NIR_API int nir_pdhg_solve(nir_ctx* ctx, const nir_lp* lp, const nir_params* p, double* x, double* y,
                           nir_info* info);

/* ---- batched solves: many LPs in persistent launches --------------------
 * No counterpart in the reference (it solves one model per call); this is
 * the batch entry point of the engine ABI. Every instance runs pdlp mode with
 * the shared parameters (mode must be NIR_MODE_PDLP; no warm start or
 * unscaling vectors) and gets the same iterates, iteration and restart counts
 * as nir_pdhg_solve on it alone. */
typedef struct nir_batch nir_batch;
typedef struct {
    int32_t threads;          /* workgroup size of the persistent kernel: 256, 512 or 1024;
                                 0 = autotune (timed on the batch after setup) */
    int32_t quantum;          /* evaluation blocks an instance runs per work-list ticket (0 = 1024:
                                 most instances finish on their first ticket; should_stop is
                                 polled between launches, so a smaller quantum stops sooner) */
    int32_t rounds_per_graph; /* persistent launches captured per hipGraph replay (0 = 4) */
    int32_t workgroups;       /* persistent grid (0 = what fits on the device) */
} nir_batch_opts;
typedef struct {
    int32_t count, threads, workgroups, quantum;
    int64_t launches;         /* persistent launches of the iteration kernel */
    int64_t total_iterations; /* summed over instances */
    double nnz_iterations;    /* sum over instances of iterations x nnz */
    double total_nnz;
    double setup_seconds, solve_seconds;
} nir_batch_stats;
NIR_API void nir_batch_opts_default(nir_batch_opts* o);
/* Uploads and prepares every instance (scaling, power iteration, start point
 * evaluation) on the device; `lps` holds `count` pointers. Each instance must
 * have fewer than 2^28 rows, columns and nonzeros (-1 otherwise). */
NIR_API int nir_batch_create(nir_ctx* ctx, const nir_lp* const* lps, int32_t count, const nir_params* p,
                             const nir_batch_opts* o, nir_batch** out);
/* The same with per-instance warm starts: x0s / y0s hold `count` pointers
 * (ncol / nrow values each; a NULL array or entry = cold start), read with
 * the single-LP meaning of nir_params.x0 / y0. */
NIR_API int nir_batch_create_warm(nir_ctx* ctx, const nir_lp* const* lps, int32_t count,
                                  const double* const* x0s, const double* const* y0s, const nir_params* p,
                                  const nir_batch_opts* o, nir_batch** out);
/* Runs every instance to its tolerance, iteration cap or should_stop (polled
 * between graph replays; unfinished instances then report status 4). */
NIR_API int nir_batch_solve(nir_batch* b);
NIR_API int nir_batch_result(nir_batch* b, int32_t i, double* x, double* y, nir_info* info);
NIR_API int nir_batch_get_stats(nir_batch* b, nir_batch_stats* out);
NIR_API void nir_batch_destroy(nir_batch* b);

/* ---- parametric families: one matrix, many right-hand sides ------------
 * No counterpart in the reference (src/gpu/pdhg_solver.hpp:25-107 solves one
 * LP; its products are the single-vector src/core/csc_matrix.hpp:99-120).
 * Members of a family share the constraint matrix and column bounds of
 * `base` and have their own costs and row bounds. The engine keeps 64
 * members per wave64 in interleaved vectors, so the matrix is streamed once
 * per 64 members and every operand gather is a coalesced 512-byte read;
 * `count` members are streamed through a pool of `lanes` lanes (finished
 * members are handed out and their lanes refilled between evaluation
 * blocks). Every member gets exactly what nir_pdhg_solve returns for it
 * alone (pdlp mode, FP64 or mixed precision): the same x, y, iteration and
 * restart counts, status and objective. */
typedef struct nir_family nir_family;
typedef struct {
    int32_t lanes;   /* members iterated together, a multiple of 64 (default 1024) */
    int32_t handoff; /* once no member waits for a lane and at most this many still run
                        (default 8; 0 = never), each of them finishes alone on the
                        single-LP engine, continued from its lane's state bit for bit:
                        a nearly empty 64-lane group iterates far slower per member */
} nir_family_opts;
typedef struct {
    int32_t lanes;
    int32_t stopped;          /* should_stop ended the last solve */
    int64_t instances;        /* members finished */
    int64_t blocks;           /* evaluation blocks run (eval_freq iterations of every lane each) */
    int64_t refills;          /* start-point evaluations of refilled lanes */
    int64_t total_iterations; /* summed over members */
    double setup_seconds, solve_seconds;
    int64_t compactions;      /* moves of running members into the lowest groups */
    int64_t group_blocks;     /* evaluation blocks x groups launched */
    double prepare_seconds;   /* host: member set-up (fill callback + scaling) */
    double wait_seconds;      /* host waiting for blocks to finish */
    int64_t handoffs;         /* members finished on the single-LP engine */
    double handoff_seconds;   /* time spent in those single-LP runs */
    int64_t tails;            /* mixed precision: members whose FP32 phase ended needing
                                 the FP64 tail (stalled, or a polish tolerance), which
                                 they ran on the single-LP engine from their state */
} nir_family_stats;
/* Supplies member `index`: cost (ncol), rowlo and rowup (nrow), and an
 * optional warm start x0 (ncol) / y0 (nrow) flagged in *warm (bit 0: x0,
 * bit 1: y0; the buffers arrive zeroed) with an optional initial primal
 * weight (*primal_weight > 0 replaces nir_params.primal_weight for this
 * member). Returns 0, or nonzero to abort. */
typedef int (*nir_family_fill_fn)(void* user, int64_t index, double* cost, double* rowlo, double* rowup,
                                  double* x0, double* y0, int32_t* warm, double* primal_weight);
/* Receives member `index`'s solution (x: ncol, y: nrow, valid during the
 * call) and its nir_info; members arrive in the order they finish. */
typedef int (*nir_family_done_fn)(void* user, int64_t index, const double* x, const double* y,
                                  const nir_info* info);
NIR_API void nir_family_opts_default(nir_family_opts* o);
/* Uploads and scales the shared matrix (pdlp mode parameters, precision FP64
 * or mixed; warm starts come per member through the fill callback). */
NIR_API int nir_family_create(nir_ctx* ctx, const nir_lp* base, const nir_params* p, const nir_family_opts* o,
                              nir_family** out);
/* Solves members 0 .. count-1; should_stop (polled between blocks) ends the
 * running members with status 4 and leaves the rest unsolved. */
NIR_API int nir_family_solve(nir_family* f, int64_t count, nir_family_fill_fn fill, nir_family_done_fn done,
                             void* user);
NIR_API int nir_family_get_stats(nir_family* f, nir_family_stats* out);
NIR_API void nir_family_destroy(nir_family* f);

/* ---- learned warm start -------------------------------------------------
 * No counterpart in the reference. A small MLP (n_in -> hidden -> hidden ->
 * n_out, GELU, bf16 MFMA GEMMs with FP32 accumulation, Adam) trained on the
 * device on solutions the engine produced, and the FP64 kernels of the
 * proper orthogonal decomposition of solution snapshots. Widths are
 * multiples of 64; parameters are FP32 in the order W1 [hidden][n_in], b1,
 * W2 [hidden][hidden], b2, W3 [n_out][hidden], b3. */
typedef struct nir_mlp nir_mlp;
NIR_API int nir_mlp_create(nir_ctx* ctx, int32_t n_in, int32_t hidden, int32_t n_out, int32_t max_rows,
                           uint64_t seed, nir_mlp** out);
NIR_API int64_t nir_mlp_num_params(const nir_mlp* m);
NIR_API int nir_mlp_get_params(nir_mlp* m, float* out);
NIR_API int nir_mlp_set_params(nir_mlp* m, const float* in);
/* out[rows][n_out] = MLP(x[rows][n_in]). */
NIR_API int nir_mlp_forward(nir_mlp* m, const float* x, int32_t rows, float* out);
/* Mean squared error over rows x n_out against target and its gradient with
 * respect to every parameter (rows <= max_rows). */
NIR_API int nir_mlp_gradient(nir_mlp* m, const float* x, const float* target, int32_t rows, float* grad,
                             double* loss);
NIR_API int nir_mlp_set_data(nir_mlp* m, const float* x, const float* target, int32_t rows);
NIR_API int nir_mlp_train(nir_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* epoch_loss);
NIR_API int nir_mlp_train_stats(nir_mlp* m, int64_t* rows, double* ms, double* gemm_flops);
NIR_API void nir_mlp_destroy(nir_mlp* m);
/* One device GEMM C[M][N] = A B^T of bf16-rounded inputs (A [M][K] or, with
 * transpose_a, [K][M]; B [N][K] or [K][N]); M, N multiples of 64, K of 32. */
NIR_API int nir_mlp_gemm_check(nir_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, const float* B,
                               int32_t transpose_a, int32_t transpose_b, float* C);
/* POD: G = X^T X and Y = X V for snapshots X [len][N] (row-major). */
NIR_API int nir_pod_gram(nir_ctx* ctx, const double* X, int32_t len, int32_t N, double* G);
NIR_API int nir_pod_project(nir_ctx* ctx, const double* X, int32_t len, int32_t N, const double* V, int32_t R,
                            double* Y);
/* A POD basis kept on the device, the source of learned start points; it
 * replaces the reference's fixed start (x from the column bounds, y = 0,
 * src/gpu/pdhg_solver.hpp:47-57) for family members. Modes [len][R] = X V
 * (X [len][N] centred snapshots, V [N][R]) are computed and kept on the
 * device with the mean [len]. */
typedef struct nir_pod nir_pod;
NIR_API int nir_pod_create(nir_ctx* ctx, const double* X, int32_t len, int32_t N, const double* V, int32_t R,
                           const double* mean, nir_pod** out);
/* Copies the modes [len][R] to the host. */
NIR_API int nir_pod_modes(nir_pod* p, double* modes);
/* rows [count][len] = mean + alpha [count][R] modes^T on the device. The rows
 * stay on the device until the next decode or destroy (the family engine's
 * device starts read them); out, when not null, receives a copy. */
NIR_API int nir_pod_decode(nir_pod* p, const double* alpha, int32_t count, double* out);
/* coef [count][R] = modes^T (s - mean) for snapshots s [count][len]. */
NIR_API int nir_pod_encode(nir_pod* p, const double* s, int32_t count, double* coef);
NIR_API void nir_pod_destroy(nir_pod* p);
/* Warm starts read on the device: in the next nir_family_solve calls member
 * i < rows starts from row i of px's last decode (x0, ncol entries) and of
 * py's (y0, nrow entries) - scaled and projected exactly as a host x0/y0 -
 * in place of the fill callback's x0/y0 (its primal weight still applies).
 * Either pod may be null; both null clears. */
NIR_API int nir_family_set_device_starts(nir_family* f, const nir_pod* px, const nir_pod* py);

/* ---- double-double solution certifier on the device -----------------------
 * Replaces the host pass of sov::verify_solution (src/io/sol_io.hpp:25-106)
 * for a continuous LP and Model::eval_objective (src/model/model.hpp:73-98):
 * relative bound and row violations (row activities in double-double,
 * ascending columns, zeros skipped - the host checker's values bit for bit),
 * the double-double objective and its relative error against claimed_obj.
 * passed = no non-finite x, both violations <= tol, objective error <= 1e-5.
 * Worst indices are the first index attaining the maximum (-1 if none). */
typedef struct {
    double max_bound_viol;
    double max_row_viol;
    double computed_obj;
    double obj_rel_err;
    int32_t worst_col;       /* column of max_bound_viol (-1 if 0) */
    int32_t worst_col_upper; /* 1 if it is the upper bound */
    int32_t worst_row;       /* row of max_row_viol (-1 if 0) */
    int32_t worst_row_upper; /* 1 if it is the row's upper bound */
    int32_t nonfinite_col;   /* first column with a NaN/Inf value, -1 if none */
    int32_t passed;
} nir_verify_report;
NIR_API int nir_verify_solution(nir_ctx* ctx, const nir_lp* lp, const double* x, double claimed_obj, double tol,
                                nir_verify_report* out);

/* Allocation-free variant of nir_verify_solution, for callers that cannot
 * call hipMalloc (tight serving loops, code that must not allocate). It has
 * no counterpart in the reference. The LP and x are already on the device, as
 * CSR (row_start int32[nrow + 1], col_index int32[nnz] ascending within each
 * row, value double[nnz]) plus the bound and cost vectors; the same kernels
 * as nir_verify_solution compute the same report bit for bit. Block partials
 * live in a static device array of 3 MiB per GPU (held until the process
 * ends), so nrow and ncol must be at most 16,777,216 (-1 otherwise) and the
 * entry point calls no hipMalloc / hipFree. Every call on one GPU shares that
 * array: calls that overlap on one GPU (two contexts, two threads) race. */
typedef struct {
    int32_t nrow;
    int32_t ncol;
    const int32_t* row_start; /* device, nrow + 1 */
    const int32_t* col_index; /* device, nnz */
    const double* value;      /* device, nnz */
    const double* cost;       /* device, ncol */
    const double* collo;      /* device, ncol */
    const double* colup;      /* device, ncol */
    const double* rowlo;      /* device, nrow */
    const double* rowup;      /* device, nrow */
    double obj_offset;
} nir_dev_lp;
NIR_API int nir_verify_solution_noalloc(nir_ctx* ctx, const nir_dev_lp* lp, const double* x_dev,
                                        double claimed_obj, double tol, nir_verify_report* out);

/* ---- presolve (host) ------------------------------------------------------
 * The facade's presolve (src/presolve/presolve.hpp:32-345, rules R1-R7, R14
 * and R2 in the reference's order and tolerances) on one LP, for callers that
 * run the engine without the facade and for parity checks. `integer` (ncol
 * flags, may be NULL) marks integer columns (rule R7 tightens their bounds).
 * Outputs in the caller's buffers, sized by the original model: fixed_x
 * (ncol, values of removed columns), col_map (reduced ncol entries: original
 * column of each kept column), row_map (reduced nrow), and the reduced
 * model's bounds collo/colup (reduced ncol) and rowlo/rowup (reduced nrow);
 * any buffer may be NULL. Returns 0, or -1 with nir_last_error() set. */
typedef struct {
    int32_t status;        /* sov Status: 0 (solve the reduced model, or solved), 1 infeasible, 2 unbounded */
    int32_t solved;        /* presolve settled the model */
    int32_t nrow, ncol;    /* reduced model */
    int32_t rows_removed, cols_removed, tightenings;
    int32_t reserved;
    double obj_offset;     /* reduced model's objective offset */
} nir_presolve_info;
NIR_API int nir_lp_presolve(const nir_lp* lp, const int8_t* integer, int32_t level, nir_presolve_info* info,
                            double* fixed_x, int32_t* col_map, int32_t* row_map, double* collo, double* colup,
                            double* rowlo, double* rowup);

/* ---- crossover hook (host) ------------------------------------------------
 * A simplex starting basis guessed from a PDHG point x (ncol), in the layout
 * of the reference's BasisSnapshot (src/lp/lp_solver.hpp:18-21) as its
 * LpSolver::set_basis takes it (:896-907): status[ncol + nrow] (the columns,
 * then one logical per row, valued at the row activity and bounded by the
 * row bounds; codes as src/core/types.hpp:40-47) and basic_index[nrow]
 * (ascending variable numbers). Variables strictly inside their bounds
 * (farther than tol * (1 + |bound|) from each finite bound) are basic, the
 * most interior m of them when there are more; logicals of the rows with the
 * most slack fill the basis up to nrow. No counterpart in the reference's
 * PDHG, which has no crossover. */
enum { NIR_VS_BASIC = 0, NIR_VS_AT_LOWER = 1, NIR_VS_AT_UPPER = 2, NIR_VS_FIXED = 3, NIR_VS_FREE_ZERO = 4 };
NIR_API int nir_basis_guess(const nir_lp* lp, const double* x, double tol, int8_t* status, int32_t* basic_index);

/* ---- synthetic instances (seeded; generated on the host) ---------------- */
typedef struct nir_owned_lp nir_owned_lp;

/* m equality rows, n columns with `per_col` distinct uniform rows each, N(0,1)
 * values, bounds [0, 10], b = A x0 with x0 ~ U[0,1], c = A^T y0 + s. */
NIR_API nir_owned_lp* nir_gen_equality_lp(uint64_t seed, int32_t m, int32_t n, int32_t per_col);
/* m rows (first half equality, second half <=), n columns with degree
 * 2 + Poisson(mean) rows, N(0,1) values, a random `bad_frac` of the columns
 * multiplied by `bad_scale`, bounds [0, U[1,100]]. */
NIR_API nir_owned_lp* nir_gen_mixed_lp(uint64_t seed, int32_t m, int32_t n, double mean, double bad_frac,
                                       double bad_scale);
/* Multi-period blending/planning LP with T periods and K products. */
NIR_API nir_owned_lp* nir_gen_staircase_lp(uint64_t seed, int32_t T, int32_t K);
/* Periods T ~ U{4..48} and products K ~ U{5..40} of the planning-family
 * member with this seed (the batch of uneven sizes). */
NIR_API void nir_gen_planning_dims(uint64_t seed, int32_t* T, int32_t* K);
NIR_API int nir_owned_lp_view(const nir_owned_lp* g, nir_lp* out);
NIR_API int64_t nir_owned_lp_nnz(const nir_owned_lp* g);
/* The feasible point used to build the right-hand side (ncol), and the dual
 * point used to build the costs (nrow). */
NIR_API const double* nir_owned_lp_x0(const nir_owned_lp* g);
NIR_API const double* nir_owned_lp_y0(const nir_owned_lp* g);
NIR_API void nir_owned_lp_free(nir_owned_lp* g);

/* Binary LP container (.nlpb, the format of sov_write_lpb) for engine-level
 * LPs: write a canonical CSC LP (min form, continuous), or read one back as an
 * owned LP (a maximisation is returned in min form; quadratic or integer
 * models are rejected). Errors: -1 / NULL with nir_last_error() set. No
 * counterpart in the reference. */
NIR_API int nir_lp_write_lpb(const nir_lp* lp, const char* path);
NIR_API nir_owned_lp* nir_lp_read_lpb(const char* path);

#ifdef __cplusplus
}
#endif
#endif
