// Parametric-family engine: many LPs that share one constraint matrix A and
// its column bounds and differ in costs and row bounds (the family
// b(theta), c(theta) of a learned warm start), solved by the restarted PDHG
// of pdlp mode in FP64 or in re-centred mixed precision, every instance on
// exactly the single-LP trajectory of that precision.
//
// Layout: instances occupy lanes. A group of 64 lanes (one wave64) keeps its
// 64 instances' vectors interleaved, element j of lane q at
// [group][j][64] + q, so a nonzero a_ij of a row product is loaded once
// (a wave-uniform scalar load) and its gathered operand entries for all 64
// instances are one contiguous 512-byte read. Each lane sums its row in
// ascending column order from 0.0, the single-LP engine's sequential order,
// so every iterate is the same bit for bit; the evaluation terms use the
// same functions (pdlp_core.hpp) and double-double sums.
//
// The matrix is streamed once per group and product instead of once per
// instance, and the gathers fetch whole cache lines of useful data: the
// random 8-byte gathers that bound the single-LP products become coalesced
// 512-byte ones.
//
// Work queue: the host streams any number of instances through a fixed pool
// of lanes. Between evaluation blocks (eval_freq iterations, one hipGraph
// replay) it harvests finished lanes (their T(z), unscaled on the host) and
// refills free lanes with the next instances: their scaled vectors and state
// are prepared on the host by the single-LP engine's own set-up code
// (pdlp_prepare_instance), scattered into the lanes, and evaluated at their
// start point (mode 1) before they iterate. Finished and empty lanes are
// skipped by every kernel (whole waves return when all 64 lanes are idle).
//
// Mixed precision (the single-LP engine's FP32 phase, pdlp.hip): each lane
// iterates on FP32 deltas around its FP64 anchor (xa / ya), with FP32
// constants re-based on it (rc = c - A^T Y, [l - X, u - X], [rl - A X,
// ru - A X], sa = -Y / sigma) and FP32 matrix values; evaluations and
// decisions are FP64 (T(z) = anchor + delta, A^T T(z)_y = A^T Y + A^T dyp). At
// a restart the anchor moves to T(z), A X and A^T Y are recomputed in FP64 and
// the constants re-based. A lane whose FP32 phase ends needing the FP64 tail
// (stalled, or a polish tolerance) continues on the single-LP engine from its
// state, which switches to FP64 exactly as it does for one LP.
//
// The reference solves one LP per call with a CPU loop
// (src/gpu/pdhg_solver.hpp:25-107, single-vector products
// src/core/csc_matrix.hpp:99-120); this engine has no counterpart there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/nirnaya/pdhg.h"
#include "ctx.hpp"
#include "devmat.hpp"
#include "lpcheck.hpp"
#include "pdlp.hpp"
#include "pdlp_core.hpp"
#include "pod.hpp"
#include "trace.hpp"

namespace nir {
namespace gpu {
namespace fam {

constexpr int kL = 64;              // lanes per group (one wave64)
constexpr int kRowsPerWave = 4;     // rows per wave in the iteration kernels
constexpr int kEvalBlocks = 256;    // evaluation workgroups per group (grid-stride over rows)
constexpr int kStageMax = 256;      // lanes refilled or harvested per transfer

struct View {
    int m, n, G;
    int ebc, ebr;   // evaluation workgroups per group over columns / rows
    const int* arp;
    const int* acol;
    const double* aval;   // CSR of the scaled A (m rows)
    const int* trp;
    const int* tcol;
    const double* tval;   // CSR of the scaled A^T (n rows)
    const double* lo;
    const double* up;        // scaled column bounds, shared
    const double* rscale;    // caller-unit scales R and C
    const double* cscale;
    double *x, *xbar, *xa, *xp, *c;    // [G][n][kL]
    double *y, *ya, *yp, *rl, *ru;     // [G][m][kL]
    double *qx, *rx;                   // certificate: old anchor, anchor step (n)
    double *qy, *ry;                   // (m)
    PdlpState* st;                     // [G * kL]
    double* ltau;
    double* lsigma;
    int* lkb;
    int* ldone;
    int* lrestart;
    int* lfresh;
    dd* part_c;   // [G][ebc][5][kL]
    dd* part_r;   // [G][ebr][6][kL]
    dd* cert_c;   // [G][ebc][6][kL]
    dd* cert_r;   // [G][ebr][6][kL]
    // mixed precision: FP32 deltas around the FP64 anchor xa / ya
    const float* aval32;                  // FP32 copies of the scaled values
    const float* tval32;
    float *fx, *fxbar, *fxp, *rc, *dl, *du;   // [G][n][kL]
    float *fy, *fyp, *drl, *dru, *sa;         // [G][m][kL]
    double* HA;                               // A X of the anchor [G][m][kL]
    double* GA;                               // A^T Y of the anchor [G][n][kL]
};

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Row r of a CSR matrix times the lane's interleaved operand v (already
// offset to the lane): sum in ascending column order from 0.0 in the
// precision T of the values, operand entries converted to T, products
// rounded before the add (no contraction), four operand reads in flight -
// the single-LP stream_rows sum for the same T and operand type.
template <typename T, typename V>
__device__ __forceinline__ T row_sum(const int* __restrict__ rp, const int* __restrict__ col,
                                     const T* __restrict__ val, const V* __restrict__ v, int r) {
    const int p0 = rp[r], p1 = rp[r + 1];
    T s = T(0);
    int p = p0;
    for (; p + 4 <= p1; p += 4) {
        const int c0 = col[p], c1 = col[p + 1], c2 = col[p + 2], c3 = col[p + 3];
        const T a0 = val[p], a1 = val[p + 1], a2 = val[p + 2], a3 = val[p + 3];
        const T g0 = static_cast<T>(v[static_cast<size_t>(c0) * kL]), g1 = static_cast<T>(v[static_cast<size_t>(c1) * kL]);
        const T g2 = static_cast<T>(v[static_cast<size_t>(c2) * kL]), g3 = static_cast<T>(v[static_cast<size_t>(c3) * kL]);
        s += a0 * g0;
        s += a1 * g1;
        s += a2 * g2;
        s += a3 * g3;
    }
    for (; p < p1; ++p) s += val[p] * static_cast<T>(v[static_cast<size_t>(col[p]) * kL]);
    return s;
}

// The same product in double-double (certificates).
__device__ __forceinline__ dd row_dot_lane(const int* __restrict__ rp, const int* __restrict__ col,
                                           const double* __restrict__ val, const double* __restrict__ v, int r) {
    dd acc = dd_zero();
    for (int p = rp[r]; p < rp[r + 1]; ++p) acc = dd_add(acc, dd_prod(val[p], v[static_cast<size_t>(col[p]) * kL]));
    return acc;
}

// Primal half of an iteration over rows of A^T: T(z)_x, the reflected point
// and the Halpern step, per lane (pdlp.hip k_primal_pdlp).
__global__ __launch_bounds__(256) void k_fam_primal(View V, int k_lo, int save) {
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    const bool done = V.ldone[li] != 0;
    if (__all(done)) return;
    const double tau = V.ltau[li];
    const Halpern<double> hw = halpern<double>(V.lkb[li] + k_lo);
    const int r0 = (blockIdx.x * 4 + wave_id()) * kRowsPerWave;
    const double* yv = V.y + static_cast<size_t>(g) * V.m * kL + lane;
    const size_t xb = static_cast<size_t>(g) * V.n * kL + lane;
    const int r1 = min(r0 + kRowsPerWave, V.n);
    if (done) return;   // finished lanes issue no loads; their waves carry on for the others
    for (int r = r0; r < r1; ++r) {
        const double gs = row_sum<double>(V.trp, V.tcol, V.tval, yv, r);
        const size_t e = xb + static_cast<size_t>(r) * kL;
        const PrimalOut<double> o = primal_point<double>(V.x[e], gs, V.c[e], V.lo[r], V.up[r], V.xa[e], tau, hw);
        V.xbar[e] = o.xb;
        V.x[e] = o.next;
        if (save) V.xp[e] = o.cand;
    }
}

// Dual half over rows of A (pdlp.hip k_dual_pdlp).
__global__ __launch_bounds__(256) void k_fam_dual(View V, int k_lo, int save) {
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    const bool done = V.ldone[li] != 0;
    if (__all(done)) return;
    const double sigma = V.lsigma[li];
    const Halpern<double> hw = halpern<double>(V.lkb[li] + k_lo);
    const int r0 = (blockIdx.x * 4 + wave_id()) * kRowsPerWave;
    const double* xv = V.xbar + static_cast<size_t>(g) * V.n * kL + lane;
    const size_t yb = static_cast<size_t>(g) * V.m * kL + lane;
    const int r1 = min(r0 + kRowsPerWave, V.m);
    if (done) return;
    for (int r = r0; r < r1; ++r) {
        const double h = row_sum<double>(V.arp, V.acol, V.aval, xv, r);
        const size_t e = yb + static_cast<size_t>(r) * kL;
        const DualOut<double> o = dual_point<double>(V.y[e], h, V.rl[e], V.ru[e], V.ya[e], sigma, hw);
        V.y[e] = o.next;
        if (save) V.yp[e] = o.yp;
    }
}

// FP32 phase of mixed precision (pdlp.hip k_primal_pdlp<float>): deltas
// around the anchor, re-based constants, a zero Halpern anchor.
__global__ __launch_bounds__(256) void k_fam_primal32(View V, int k_lo, int save) {
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    const bool done = V.ldone[li] != 0;
    if (__all(done)) return;
    const float tau = static_cast<float>(V.ltau[li]);
    const Halpern<float> hw = halpern<float>(V.lkb[li] + k_lo);
    const int r0 = (blockIdx.x * 4 + wave_id()) * kRowsPerWave;
    const float* yv = V.fy + static_cast<size_t>(g) * V.m * kL + lane;
    const size_t xb = static_cast<size_t>(g) * V.n * kL + lane;
    const int r1 = min(r0 + kRowsPerWave, V.n);
    if (done) return;
    for (int r = r0; r < r1; ++r) {
        const float gs = row_sum<float>(V.trp, V.tcol, V.tval32, yv, r);
        const size_t e = xb + static_cast<size_t>(r) * kL;
        const PrimalOut<float> o = primal_point<float>(V.fx[e], gs, V.rc[e], V.dl[e], V.du[e], 0.0f, tau, hw);
        V.fxbar[e] = o.xb;
        V.fx[e] = o.next;
        if (save) V.fxp[e] = o.cand;
    }
}

// pdlp.hip k_dual_pdlp<float>: sa = -Y / sigma stands in the anchor's place.
__global__ __launch_bounds__(256) void k_fam_dual32(View V, int k_lo, int save) {
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    const bool done = V.ldone[li] != 0;
    if (__all(done)) return;
    const float sigma = static_cast<float>(V.lsigma[li]);
    const Halpern<float> hw = halpern<float>(V.lkb[li] + k_lo);
    const int r0 = (blockIdx.x * 4 + wave_id()) * kRowsPerWave;
    const float* xv = V.fxbar + static_cast<size_t>(g) * V.n * kL + lane;
    const size_t yb = static_cast<size_t>(g) * V.m * kL + lane;
    const int r1 = min(r0 + kRowsPerWave, V.m);
    if (done) return;
    for (int r = r0; r < r1; ++r) {
        const float h = row_sum<float>(V.arp, V.acol, V.aval32, xv, r);
        const size_t e = yb + static_cast<size_t>(r) * kL;
        const DualOut<float> o = dual_point_delta<float>(V.fy[e], h, V.drl[e], V.dru[e], V.sa[e], sigma, hw);
        V.fy[e] = o.next;
        if (save) V.fyp[e] = o.yp;
    }
}

// Lanes an evaluation pass serves: fresh lanes (start point, mode 1) or the
// running ones (after a block, mode 0).
template <int FRESH>
__device__ __forceinline__ bool eval_lane(const View& V, int li) {
    if constexpr (FRESH) return V.lfresh[li] != 0;
    else return V.ldone[li] == 0;
}

// Sum NQ dd terms per lane over the workgroup's four waves; wave 0 writes
// them to part[((g * blocks + b) * NQ + q) * kL + lane].
template <int NQ>
__device__ __forceinline__ void waves_to_part(dd (&acc)[NQ], dd* sh, dd* part, int g, int blocks) {
    const int lane = threadIdx.x & (kL - 1), w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) sh[(w * NQ + q) * kL + lane] = acc[q];
    __syncthreads();
    if (w == 0) {
        for (int ww = 1; ww < 4; ++ww)
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] = dd_add(acc[q], sh[(ww * NQ + q) * kL + lane]);
        dd* out = part + (static_cast<size_t>(g) * blocks + blockIdx.x) * NQ * kL + lane;
#pragma unroll
        for (int q = 0; q < NQ; ++q) out[static_cast<size_t>(q) * kL] = acc[q];
    }
}

// Column evaluation terms (pdlp.hip k_eval_cols): A^T yp per lane. FRESH
// (a template parameter, so each pass is straight-line code) selects the
// lanes: refilled ones at their start point, or the running ones. DELTA: the
// FP32 phase, T(z) = anchor + delta and A^T yp = GA + A^T dyp (FP64 values,
// FP32 operand converted).
template <int FRESH, int DELTA>
__global__ __launch_bounds__(256) void k_fam_eval_cols(View V) {
    __shared__ dd sh[4 * 5 * kL];
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    const bool serve = eval_lane<FRESH>(V, li);
    if (__all(!serve)) return;
    dd acc[5] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    const size_t yo = static_cast<size_t>(g) * V.m * kL + lane;
    const size_t xb = static_cast<size_t>(g) * V.n * kL + lane;
    for (int r = blockIdx.x * 4 + wave_id(); serve && r < V.n; r += gridDim.x * 4) {
        const size_t e = xb + static_cast<size_t>(r) * kL;
        dd t[5];
        if constexpr (DELTA) {
            const double gs = row_sum<double>(V.trp, V.tcol, V.tval, V.fyp + yo, r);
            const double X = V.xa[e];
            col_terms(V.GA[e] + gs, V.c[e], V.lo[r], V.up[r], V.cscale[r], X + static_cast<double>(V.fxp[e]), X, t);
        } else {
            const double gs = row_sum<double>(V.trp, V.tcol, V.tval, V.yp + yo, r);
            col_terms(gs, V.c[e], V.lo[r], V.up[r], V.cscale[r], V.xp[e], V.xa[e], t);
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) acc[q] = dd_add(acc[q], t[q]);
    }
    waves_to_part<5>(acc, sh, V.part_c, g, V.ebc);
}

// Row evaluation terms (pdlp.hip k_eval_rows): A xp per lane.
template <int FRESH, int DELTA>
__global__ __launch_bounds__(256) void k_fam_eval_rows(View V) {
    __shared__ dd sh[4 * 6 * kL];
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    const bool serve = eval_lane<FRESH>(V, li);
    if (__all(!serve)) return;
    dd acc[6] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    const size_t xo = static_cast<size_t>(g) * V.n * kL + lane;
    const size_t yb = static_cast<size_t>(g) * V.m * kL + lane;
    for (int r = blockIdx.x * 4 + wave_id(); serve && r < V.m; r += gridDim.x * 4) {
        const size_t e = yb + static_cast<size_t>(r) * kL;
        dd t[6];
        if constexpr (DELTA) {
            const double h = row_sum<double>(V.arp, V.acol, V.aval, V.fxp + xo, r);
            const double Y = V.ya[e];
            row_terms(V.HA[e] + h, V.rl[e], V.ru[e], V.rscale[r], Y + static_cast<double>(V.fyp[e]), Y, t);
        } else {
            const double h = row_sum<double>(V.arp, V.acol, V.aval, V.xp + xo, r);
            row_terms(h, V.rl[e], V.ru[e], V.rscale[r], V.yp[e], V.ya[e], t);
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[q] = dd_add(acc[q], t[q]);
    }
    waves_to_part<6>(acc, sh, V.part_r, g, V.ebr);
}

// Per-lane reduction of the evaluation partials and the decision
// (pdlp_decide, as k_decide). One workgroup of 64 lanes per group.
template <int FRESH>
__global__ __launch_bounds__(kL) void k_fam_decide(View V) {
    const int g = blockIdx.x, lane = threadIdx.x, li = g * kL + lane;
    if (!eval_lane<FRESH>(V, li)) return;
    dd tot[11];
    for (int q = 0; q < 11; ++q) tot[q] = dd_zero();
    for (int b = 0; b < V.ebc; ++b) {
        const dd* p = V.part_c + (static_cast<size_t>(g) * V.ebc + b) * 5 * kL + lane;
        for (int q = 0; q < 5; ++q) tot[q] = dd_add(tot[q], p[static_cast<size_t>(q) * kL]);
    }
    for (int b = 0; b < V.ebr; ++b) {
        const dd* p = V.part_r + (static_cast<size_t>(g) * V.ebr + b) * 6 * kL + lane;
        for (int q = 0; q < 6; ++q) tot[5 + q] = dd_add(tot[5 + q], p[static_cast<size_t>(q) * kL]);
    }
    PdlpState s = V.st[li];
    pdlp_decide(s, tot, FRESH);
    V.st[li] = s;
    V.ltau[li] = s.tau;
    V.lsigma[li] = s.sigma;
    V.lkb[li] = s.k_base;
    V.ldone[li] = s.done;
    V.lrestart[li] = s.restart;
    if constexpr (FRESH) V.lfresh[li] = 0;
}

__device__ __forceinline__ bool cert_lane(const View& V, int li) { return !V.ldone[li] && V.lrestart[li]; }

// Restart bookkeeping over the interleaved vectors; blockIdx.y = group.
// STAGE 0: keep the old anchor (k_cert_save); 1: iterate and anchor := T(z)
// (k_restart_copy); 2: the anchor step (k_cert_diff); 3: the FP32 phase's
// anchor moves to T(z) = anchor + delta (k_anchor_move). The stage is a
// template parameter, not a kernel argument: one straight-line kernel per
// stage.
template <int STAGE>
__global__ __launch_bounds__(256) void k_fam_restart(View V) {
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    if (!cert_lane(V, li)) return;
    const size_t nx = static_cast<size_t>(V.n) * kL, ny = static_cast<size_t>(V.m) * kL;
    const size_t bx = static_cast<size_t>(g) * nx, by = static_cast<size_t>(g) * ny;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t t0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t t = t0; t < nx; t += stride) {
        const size_t e = bx + t;
        if constexpr (STAGE == 0) {
            V.qx[e] = V.xa[e];
        } else if constexpr (STAGE == 1) {
            const double v = V.xp[e];
            V.x[e] = v;
            V.xa[e] = v;
        } else if constexpr (STAGE == 2) {
            V.rx[e] = V.xa[e] - V.qx[e];
        } else {
            V.xa[e] = V.xa[e] + static_cast<double>(V.fxp[e]);
        }
    }
    for (size_t t = t0; t < ny; t += stride) {
        const size_t e = by + t;
        if constexpr (STAGE == 0) {
            V.qy[e] = V.ya[e];
        } else if constexpr (STAGE == 1) {
            const double v = V.yp[e];
            V.y[e] = v;
            V.ya[e] = v;
        } else if constexpr (STAGE == 2) {
            V.ry[e] = V.ya[e] - V.qy[e];
        } else {
            V.ya[e] = V.ya[e] + static_cast<double>(V.fyp[e]);
        }
    }
}

// Lanes re-centred: refilled lanes at their start point (FORCE, as the
// single-LP set-up does) or lanes whose decision restarted.
template <int FORCE>
__device__ __forceinline__ bool recenter_lane(const View& V, int li) {
    if constexpr (FORCE) return V.lfresh[li] != 0;
    else return cert_lane(V, li);
}

// The anchor's products in FP64 (pdlp.hip k_spmv_on_restart): HA = A X over
// the rows of A, GA = A^T Y over the rows of A^T (blockIdx.x below the first
// `split` workgroups), 4 rows per wave.
template <int FORCE>
__global__ __launch_bounds__(256) void k_fam_anchor(View V, int split) {
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    const bool serve = recenter_lane<FORCE>(V, li);
    if (__all(!serve)) return;
    if (!serve) return;
    const bool rows_of_a = static_cast<int>(blockIdx.x) < split;
    const int bx = rows_of_a ? blockIdx.x : blockIdx.x - split;
    const int r0 = (bx * 4 + wave_id()) * kRowsPerWave;
    if (rows_of_a) {
        const double* xv = V.xa + static_cast<size_t>(g) * V.n * kL + lane;
        const size_t yb = static_cast<size_t>(g) * V.m * kL + lane;
        const int r1 = min(r0 + kRowsPerWave, V.m);
        for (int r = r0; r < r1; ++r) V.HA[yb + static_cast<size_t>(r) * kL] = row_sum<double>(V.arp, V.acol, V.aval, xv, r);
    } else {
        const double* yv = V.ya + static_cast<size_t>(g) * V.m * kL + lane;
        const size_t xb = static_cast<size_t>(g) * V.n * kL + lane;
        const int r1 = min(r0 + kRowsPerWave, V.n);
        for (int r = r0; r < r1; ++r) V.GA[xb + static_cast<size_t>(r) * kL] = row_sum<double>(V.trp, V.tcol, V.tval, yv, r);
    }
}

// FP32 constants re-based on the anchor and zero deltas (pdlp.hip
// k_recenter); sigma is the lane's step after its decision.
template <int FORCE>
__global__ __launch_bounds__(256) void k_fam_recenter(View V) {
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    if (!recenter_lane<FORCE>(V, li)) return;
    const double sigma = V.st[li].sigma;
    const size_t nx = static_cast<size_t>(V.n) * kL, ny = static_cast<size_t>(V.m) * kL;
    const size_t bx = static_cast<size_t>(g) * nx, by = static_cast<size_t>(g) * ny;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t t0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t t = t0; t < nx; t += stride) {
        const size_t e = bx + t;
        const size_t j = t / kL;
        const double X = V.xa[e];
        V.rc[e] = static_cast<float>(V.c[e] - V.GA[e]);
        V.dl[e] = static_cast<float>(V.lo[j] - X);
        V.du[e] = static_cast<float>(V.up[j] - X);
        V.fx[e] = 0.0f;
        V.fxp[e] = 0.0f;
    }
    for (size_t t = t0; t < ny; t += stride) {
        const size_t e = by + t;
        const double H = V.HA[e];
        V.drl[e] = static_cast<float>(V.rl[e] - H);
        V.dru[e] = static_cast<float>(V.ru[e] - H);
        V.sa[e] = static_cast<float>(-V.ya[e] / sigma);
        V.fy[e] = 0.0f;
        V.fyp[e] = 0.0f;
    }
}

// Certificate terms (pdlp.hip k_cert_cols / k_cert_rows) per lane.
__global__ __launch_bounds__(256) void k_fam_cert_cols(View V) {
    __shared__ dd sh[4 * 6 * kL];
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    const bool serve = cert_lane(V, li);
    if (__all(!serve)) return;
    dd acc[6] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    const double* dyv = V.ry + static_cast<size_t>(g) * V.m * kL + lane;
    const size_t xb = static_cast<size_t>(g) * V.n * kL + lane;
    for (int r = blockIdx.x * 4 + wave_id(); serve && r < V.n; r += gridDim.x * 4) {
        const double lam = -dd_val(row_dot_lane(V.trp, V.tcol, V.tval, dyv, r));
        const size_t e = xb + static_cast<size_t>(r) * kL;
        cert_col_terms(lam, V.c[e], V.lo[r], V.up[r], V.rx[e], acc);
    }
    waves_to_part<6>(acc, sh, V.cert_c, g, V.ebc);
}

__global__ __launch_bounds__(256) void k_fam_cert_rows(View V) {
    __shared__ dd sh[4 * 6 * kL];
    const int g = blockIdx.y;
    const int lane = threadIdx.x & (kL - 1);
    const int li = g * kL + lane;
    const bool serve = cert_lane(V, li);
    if (__all(!serve)) return;
    dd acc[6] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    const double* dxv = V.rx + static_cast<size_t>(g) * V.n * kL + lane;
    const size_t yb = static_cast<size_t>(g) * V.m * kL + lane;
    for (int r = blockIdx.x * 4 + wave_id(); serve && r < V.m; r += gridDim.x * 4) {
        const double h = dd_val(row_dot_lane(V.arp, V.acol, V.aval, dxv, r));
        const size_t e = yb + static_cast<size_t>(r) * kL;
        cert_row_terms(h, V.rl[e], V.ru[e], V.ry[e], acc);
    }
    waves_to_part<6>(acc, sh, V.cert_r, g, V.ebr);
}

__global__ __launch_bounds__(kL) void k_fam_cert_decide(View V) {
    const int g = blockIdx.x, lane = threadIdx.x, li = g * kL + lane;
    if (!cert_lane(V, li)) return;
    dd tot[6];
    for (int q = 0; q < 6; ++q) tot[q] = dd_zero();
    for (int b = 0; b < V.ebc; ++b) {
        const dd* p = V.cert_c + (static_cast<size_t>(g) * V.ebc + b) * 6 * kL + lane;
        for (int q = 0; q < 6; ++q) tot[q] = dd_add(tot[q], p[static_cast<size_t>(q) * kL]);
    }
    for (int b = 0; b < V.ebr; ++b) {
        const dd* p = V.cert_r + (static_cast<size_t>(g) * V.ebr + b) * 6 * kL + lane;
        for (int q = 0; q < 6; ++q) tot[q] = dd_add(tot[q], p[static_cast<size_t>(q) * kL]);
    }
    const int code = cert_decide(tot);
    if (code) {
        V.st[li].done = 1;
        V.st[li].status = code;
        V.st[li].restart = 0;
        V.ldone[li] = 1;
        V.lrestart[li] = 0;
    }
}

// Refill: slot f (blockIdx.y) of the staging arrays goes into lane lanes[f].
// A start taken on the device (src[f] >= 0 for x, src[count + f] for y) is
// row src of dx0 / dy0 in caller units, scaled and projected as the host
// set-up does it: x = clamp(x0 / C, lo, up), y = y0 / R.
__global__ __launch_bounds__(256) void k_fam_scatter(View V, int count, const int* __restrict__ lanes,
                                                     const double* __restrict__ sc, const double* __restrict__ sx,
                                                     const double* __restrict__ srl, const double* __restrict__ sru,
                                                     const double* __restrict__ sy, const PdlpState* __restrict__ sst,
                                                     const long long* __restrict__ src,
                                                     const double* __restrict__ dx0, const double* __restrict__ dy0) {
    const int f = blockIdx.y;
    if (f >= count) return;
    const int li = lanes[f];
    const int g = li / kL, q = li % kL;
    const size_t bx = static_cast<size_t>(g) * V.n * kL + q, by = static_cast<size_t>(g) * V.m * kL + q;
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    const long long srx = src[f], sry = src[count + f];
    for (int j = t0; j < V.n; j += stride) {
        const size_t e = bx + static_cast<size_t>(j) * kL;
        double v;
        if (srx >= 0) {
            v = dx0[static_cast<size_t>(srx) * V.n + j] / V.cscale[j];
            v = (V.lo[j] < v) ? v : V.lo[j];
            v = (v < V.up[j]) ? v : V.up[j];
        } else {
            v = sx[static_cast<size_t>(f) * V.n + j];
        }
        V.x[e] = v;
        V.xbar[e] = v;
        V.xa[e] = v;
        V.xp[e] = v;
        V.c[e] = sc[static_cast<size_t>(f) * V.n + j];
    }
    for (int i = t0; i < V.m; i += stride) {
        const size_t e = by + static_cast<size_t>(i) * kL;
        const double v = sry >= 0 ? dy0[static_cast<size_t>(sry) * V.m + i] / V.rscale[i]
                                  : sy[static_cast<size_t>(f) * V.m + i];
        V.y[e] = v;
        V.ya[e] = v;
        V.yp[e] = v;
        V.rl[e] = srl[static_cast<size_t>(f) * V.m + i];
        V.ru[e] = sru[static_cast<size_t>(f) * V.m + i];
    }
    if (t0 == 0) {
        const PdlpState s = sst[f];
        V.st[li] = s;
        V.ltau[li] = s.tau;
        V.lsigma[li] = s.sigma;
        V.lkb[li] = 0;
        V.ldone[li] = 0;
        V.lrestart[li] = 0;
        V.lfresh[li] = 1;
    }
}

// Compaction: lane src[f] moves to the free lane dst[f] (every vector an
// iteration or evaluation block reads, and the state); src is left empty.
template <int MIXED>
__global__ __launch_bounds__(256) void k_fam_move(View V, int count, const int* __restrict__ src,
                                                  const int* __restrict__ dst) {
    const int f = blockIdx.y;
    if (f >= count) return;
    const int ls = src[f], ld = dst[f];
    const int gs = ls / kL, qs = ls % kL, gd = ld / kL, qd = ld % kL;
    const size_t sx = static_cast<size_t>(gs) * V.n * kL + qs, dx = static_cast<size_t>(gd) * V.n * kL + qd;
    const size_t sy = static_cast<size_t>(gs) * V.m * kL + qs, dy = static_cast<size_t>(gd) * V.m * kL + qd;
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int j = t0; j < V.n; j += stride) {
        const size_t a = sx + static_cast<size_t>(j) * kL, b = dx + static_cast<size_t>(j) * kL;
        V.x[b] = V.x[a];
        V.xbar[b] = V.xbar[a];
        V.xa[b] = V.xa[a];
        V.xp[b] = V.xp[a];
        V.c[b] = V.c[a];
        if constexpr (MIXED) {
            V.fx[b] = V.fx[a];
            V.fxbar[b] = V.fxbar[a];
            V.fxp[b] = V.fxp[a];
            V.rc[b] = V.rc[a];
            V.dl[b] = V.dl[a];
            V.du[b] = V.du[a];
            V.GA[b] = V.GA[a];
        }
    }
    for (int i = t0; i < V.m; i += stride) {
        const size_t a = sy + static_cast<size_t>(i) * kL, b = dy + static_cast<size_t>(i) * kL;
        V.y[b] = V.y[a];
        V.ya[b] = V.ya[a];
        V.yp[b] = V.yp[a];
        V.rl[b] = V.rl[a];
        V.ru[b] = V.ru[a];
        if constexpr (MIXED) {
            V.fy[b] = V.fy[a];
            V.fyp[b] = V.fyp[a];
            V.drl[b] = V.drl[a];
            V.dru[b] = V.dru[a];
            V.sa[b] = V.sa[a];
            V.HA[b] = V.HA[a];
        }
    }
    if (t0 == 0) {
        V.st[ld] = V.st[ls];
        V.ltau[ld] = V.ltau[ls];
        V.lsigma[ld] = V.lsigma[ls];
        V.lkb[ld] = V.lkb[ls];
        V.ldone[ld] = V.ldone[ls];
        V.lrestart[ld] = V.lrestart[ls];
        V.lfresh[ld] = V.lfresh[ls];
        V.ldone[ls] = 1;
        V.lrestart[ls] = 0;
        V.lfresh[ls] = 0;
    }
}

// Harvest: T(z) and the scaled cost of lane lanes[f] into slot f (in the FP32
// phase T(z) = anchor + delta, in FP64).
template <int MIXED>
__global__ __launch_bounds__(256) void k_fam_gather(View V, int count, const int* __restrict__ lanes,
                                                    double* __restrict__ ox, double* __restrict__ oc,
                                                    double* __restrict__ oy, PdlpState* __restrict__ ost) {
    const int f = blockIdx.y;
    if (f >= count) return;
    const int li = lanes[f];
    const int g = li / kL, q = li % kL;
    const size_t bx = static_cast<size_t>(g) * V.n * kL + q, by = static_cast<size_t>(g) * V.m * kL + q;
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int j = t0; j < V.n; j += stride) {
        const size_t e = bx + static_cast<size_t>(j) * kL;
        ox[static_cast<size_t>(f) * V.n + j] = MIXED ? V.xa[e] + static_cast<double>(V.fxp[e]) : V.xp[e];
        oc[static_cast<size_t>(f) * V.n + j] = V.c[e];
    }
    for (int i = t0; i < V.m; i += stride) {
        const size_t e = by + static_cast<size_t>(i) * kL;
        oy[static_cast<size_t>(f) * V.m + i] = MIXED ? V.ya[e] + static_cast<double>(V.fyp[e]) : V.yp[e];
    }
    if (t0 == 0) ost[f] = V.st[li];
}

// Hand-off: the whole state of lane li into contiguous vectors, x, xbar, xa,
// xp, c (n each) then y, ya, yp, rl, ru (m each); with MIXED also GA (n) and
// HA (m) after those, and the FP32 vectors fx, fxbar, fxp, rc, dl, du (n
// each) then fy, fyp, drl, dru, sa (m each) into out32.
template <int MIXED>
__global__ __launch_bounds__(256) void k_fam_extract(View V, int li, double* __restrict__ out,
                                                     float* __restrict__ out32) {
    const int g = li / kL, q = li % kL;
    const size_t bx = static_cast<size_t>(g) * V.n * kL + q, by = static_cast<size_t>(g) * V.m * kL + q;
    const size_t n = V.n, m = V.m;
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int j = t0; j < V.n; j += stride) {
        const size_t e = bx + static_cast<size_t>(j) * kL;
        out[j] = V.x[e];
        out[n + j] = V.xbar[e];
        out[2 * n + j] = V.xa[e];
        out[3 * n + j] = V.xp[e];
        out[4 * n + j] = V.c[e];
        if constexpr (MIXED) {
            out[5 * n + 5 * m + j] = V.GA[e];
            out32[j] = V.fx[e];
            out32[n + j] = V.fxbar[e];
            out32[2 * n + j] = V.fxp[e];
            out32[3 * n + j] = V.rc[e];
            out32[4 * n + j] = V.dl[e];
            out32[5 * n + j] = V.du[e];
        }
    }
    double* oy = out + 5 * n;
    for (int i = t0; i < V.m; i += stride) {
        const size_t e = by + static_cast<size_t>(i) * kL;
        oy[i] = V.y[e];
        oy[m + i] = V.ya[e];
        oy[2 * m + i] = V.yp[e];
        oy[3 * m + i] = V.rl[e];
        oy[4 * m + i] = V.ru[e];
        if constexpr (MIXED) {
            out[6 * n + 5 * m + i] = V.HA[e];
            float* fy = out32 + 6 * n;
            fy[i] = V.fy[e];
            fy[m + i] = V.fyp[e];
            fy[2 * m + i] = V.drl[e];
            fy[3 * m + i] = V.dru[e];
            fy[4 * m + i] = V.sa[e];
        }
    }
}

// Marks lane li empty (after a hand-off).
__global__ void k_fam_release(View V, int li) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        V.ldone[li] = 1;
        V.lrestart[li] = 0;
        V.lfresh[li] = 0;
    }
}

// FP32 copies of the scaled matrix values (mixed precision).
__global__ void k_fam_to_float(const double* __restrict__ in, float* __restrict__ out, long long n) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = t; i < n; i += stride) out[i] = static_cast<float>(in[i]);
}

// Stop request: every running lane ends with status 4 (time limit).
__global__ void k_fam_stop(View V, int lanes) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= lanes || V.ldone[li]) return;
    V.st[li].done = 1;
    V.st[li].status = 4;
    V.ldone[li] = 1;
}

}  // namespace fam
}  // namespace gpu
}  // namespace nir

using namespace nir::gpu;
using namespace nir::gpu::fam;

struct nir_family {
    nir_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    nir_params prm{};
    nir_family_opts opts{};
    int m = 0, n = 0;
    long long nnz = 0;
    int G = 0, lanes = 0;
    int Gl = 0;   // groups launched: those up to the highest occupied lane
    // shared data
    std::vector<int64_t> col_start;
    std::vector<int32_t> row_index;
    std::vector<double> value, collo, colup;
    double offset = 0.0;
    std::vector<double> R, C;
    double eta = 0.0;
    DevMatrix A, AT;
    DevArray<double> dR, dC, lo, up;
    // lanes
    DevArray<double> x, xbar, xa, xp, c, y, ya, yp, rl, ru, qx, rx, qy, ry;
    DevArray<PdlpState> st;
    DevArray<double> ltau, lsigma;
    DevArray<int> lkb, ldone, lrestart, lfresh;
    DevArray<dd> part_c, part_r, cert_c, cert_r;
    // staging
    DevArray<int> s_lanes, s_dst;
    DevArray<double> s_c, s_x, s_rl, s_ru, s_y;
    DevArray<PdlpState> s_st;
    DevArray<long long> s_src;   // [2][S]: device start row of x / y per slot, -1 = staged
    // mixed precision
    bool mixed = false;
    DevArray<float> aval32, tval32, fx, fxbar, fxp, rc, dl, du, fy, fyp, drl, dru, sa;
    DevArray<double> HA, GA;
    std::vector<int> h_done;
    // device warm starts (nir_family_set_device_starts)
    const nir_pod* start_x = nullptr;
    const nir_pod* start_y = nullptr;
    int ebc = 1, ebr = 1;
    // graph of one evaluation block
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_groups = 0;   // Gl the graph was captured with
    nir_family_stats stats{};

    ~nir_family() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }

    View view() {
        View V{};
        V.m = m;
        V.n = n;
        V.G = G;
        V.ebc = ebc;
        V.ebr = ebr;
        V.arp = A.shape.rowptr.p;
        V.acol = A.shape.col.p;
        V.aval = A.val.p;
        V.trp = AT.shape.rowptr.p;
        V.tcol = AT.shape.col.p;
        V.tval = AT.val.p;
        V.lo = lo.p;
        V.up = up.p;
        V.rscale = dR.p;
        V.cscale = dC.p;
        V.x = x.p; V.xbar = xbar.p; V.xa = xa.p; V.xp = xp.p; V.c = c.p;
        V.y = y.p; V.ya = ya.p; V.yp = yp.p; V.rl = rl.p; V.ru = ru.p;
        V.qx = qx.p; V.rx = rx.p; V.qy = qy.p; V.ry = ry.p;
        V.st = st.p;
        V.ltau = ltau.p;
        V.lsigma = lsigma.p;
        V.lkb = lkb.p;
        V.ldone = ldone.p;
        V.lrestart = lrestart.p;
        V.lfresh = lfresh.p;
        V.part_c = part_c.p;
        V.part_r = part_r.p;
        V.cert_c = cert_c.p;
        V.cert_r = cert_r.p;
        V.aval32 = aval32.p;
        V.tval32 = tval32.p;
        V.fx = fx.p; V.fxbar = fxbar.p; V.fxp = fxp.p; V.rc = rc.p; V.dl = dl.p; V.du = du.p;
        V.fy = fy.p; V.fyp = fyp.p; V.drl = drl.p; V.dru = dru.p; V.sa = sa.p;
        V.HA = HA.p;
        V.GA = GA.p;
        return V;
    }
};

namespace {

int iter_grid(int rows) { return std::max(1, (rows + 4 * kRowsPerWave - 1) / (4 * kRowsPerWave)); }

template <int FRESH>
void launch_eval(nir_family* F, const View& V) {
    if (F->mixed) {
        k_fam_eval_cols<FRESH, 1><<<dim3(F->ebc, F->Gl), 256, 0, F->s>>>(V);
        k_fam_eval_rows<FRESH, 1><<<dim3(F->ebr, F->Gl), 256, 0, F->s>>>(V);
    } else {
        k_fam_eval_cols<FRESH, 0><<<dim3(F->ebc, F->Gl), 256, 0, F->s>>>(V);
        k_fam_eval_rows<FRESH, 0><<<dim3(F->ebr, F->Gl), 256, 0, F->s>>>(V);
    }
    k_fam_decide<FRESH><<<F->Gl, kL, 0, F->s>>>(V);
}

int restart_grid(const nir_family* F) { return std::max(1, std::min(1024, (std::max(F->n, F->m) * kL + 255) / 256)); }

// Re-centring of the FP32 phase (pdlp.hip recenter): the anchor's products
// in FP64, then the re-based constants and zero deltas. FORCE: the refilled
// lanes at their start point; else the lanes whose decision restarted, after
// their anchor moved.
template <int FORCE>
void launch_recenter(nir_family* F, const View& V) {
    const int split = iter_grid(F->m);
    k_fam_anchor<FORCE><<<dim3(split + iter_grid(F->n), F->Gl), 256, 0, F->s>>>(V, split);
    k_fam_recenter<FORCE><<<dim3(restart_grid(F), F->Gl), 256, 0, F->s>>>(V);
}

// One evaluation block: eval_freq iterations, the evaluation and decision,
// the restart copy and the certificate test of a restart (the single-LP
// order of pdlp.hip launch_eval in FP64).
void enqueue_block(nir_family* F) {
    const View V = F->view();
    const int ef = std::max(1, F->prm.eval_freq);
    const dim3 gp(iter_grid(F->n), F->Gl), gd(iter_grid(F->m), F->Gl);
    for (int kl = 0; kl < ef; ++kl) {
        const int save = kl == ef - 1 ? 1 : 0;
        if (F->mixed) {
            k_fam_primal32<<<gp, 256, 0, F->s>>>(V, kl, save);
            k_fam_dual32<<<gd, 256, 0, F->s>>>(V, kl, save);
        } else {
            k_fam_primal<<<gp, 256, 0, F->s>>>(V, kl, save);
            k_fam_dual<<<gd, 256, 0, F->s>>>(V, kl, save);
        }
    }
    launch_eval<0>(F, V);
    const int gv = restart_grid(F);
    k_fam_restart<0><<<dim3(gv, F->Gl), 256, 0, F->s>>>(V);
    if (F->mixed) {
        k_fam_restart<3><<<dim3(gv, F->Gl), 256, 0, F->s>>>(V);
        launch_recenter<0>(F, V);
    } else {
        k_fam_restart<1><<<dim3(gv, F->Gl), 256, 0, F->s>>>(V);
    }
    k_fam_restart<2><<<dim3(gv, F->Gl), 256, 0, F->s>>>(V);
    k_fam_cert_cols<<<dim3(F->ebc, F->Gl), 256, 0, F->s>>>(V);
    k_fam_cert_rows<<<dim3(F->ebr, F->Gl), 256, 0, F->s>>>(V);
    k_fam_cert_decide<<<F->Gl, kL, 0, F->s>>>(V);
    HIP_CHECK(hipGetLastError());
}

void run_block(nir_family* F) {
    if (F->prm.use_graph) {
        if (F->gexec && F->graph_groups != F->Gl) {
            // Compaction changed the launched groups: capture again.
            HIP_CHECK(hipGraphExecDestroy(F->gexec));
            HIP_CHECK(hipGraphDestroy(F->graph));
            F->gexec = nullptr;
            F->graph = nullptr;
        }
        if (!F->gexec) {
            HIP_CHECK(hipStreamBeginCapture(F->s, hipStreamCaptureModeThreadLocal));
            enqueue_block(F);
            HIP_CHECK(hipStreamEndCapture(F->s, &F->graph));
            HIP_CHECK(hipGraphInstantiate(&F->gexec, F->graph, nullptr, nullptr, 0));
            F->graph_groups = F->Gl;
        }
        HIP_CHECK(hipGraphLaunch(F->gexec, F->s));
    } else {
        enqueue_block(F);
    }
}

// A prepared instance waiting for a lane.
struct Prepared {
    long long index;
    PdlpInstance inst;
    bool dev_x = false, dev_y = false;   // start rows read on the device
};

}  // namespace

extern "C" {

NIR_API void nir_family_opts_default(nir_family_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->lanes = 1024;
    o->handoff = 8;
}

NIR_API int nir_family_create(nir_ctx* ctx, const nir_lp* base, const nir_params* p, const nir_family_opts* o,
                              nir_family** out) {
    const nir::TraceRange trace("nir_family_create");
    if (!ctx || !out) return -1;
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto F = std::make_unique<nir_family>();
    try {
        check_lp(base);
        if (!p) throw std::invalid_argument("family needs parameters");
        F->prm = *p;
        if (o) F->opts = *o;
        else nir_family_opts_default(&F->opts);
        if (F->prm.mode != NIR_MODE_PDLP) throw std::invalid_argument("the family engine runs pdlp mode only");
        if (F->prm.precision != NIR_PREC_FP64 && F->prm.precision != NIR_PREC_MIXED)
            throw std::invalid_argument("unknown precision");
        F->mixed = F->prm.precision == NIR_PREC_MIXED;
        if (F->prm.x0 || F->prm.y0 || F->prm.unscale_row || F->prm.unscale_col)
            throw std::invalid_argument("family solves take warm starts per instance and no unscaling vectors");
        if (F->opts.lanes <= 0 || F->opts.lanes % kL != 0)
            throw std::invalid_argument("family lanes must be a positive multiple of 64");
        HIP_CHECK(hipSetDevice(ctx->device));
        F->ctx = ctx;
        F->s = ctx->stream;
        const int m = base->nrow, n = base->ncol;
        F->m = m;
        F->n = n;
        F->nnz = n > 0 ? base->col_start[n] : 0;
        if (F->nnz >= (1LL << 31)) throw std::invalid_argument("family matrix has 2^31 or more nonzeros");
        // Each lane sums a row in sequential order, which is pdlp mode's order
        // up to kLongRow nonzeros only.
        if (longest_line(base) > kLongRow)
            throw std::invalid_argument("family matrices with a row or column of more than 4096 nonzeros are not "
                                        "supported: solve the members with the single-LP engine");
        F->col_start.assign(base->col_start, base->col_start + n + 1);
        F->row_index.assign(base->row_index, base->row_index + F->nnz);
        F->value.assign(base->value, base->value + F->nnz);
        F->collo.assign(base->collo, base->collo + n);
        F->colup.assign(base->colup, base->colup + n);
        F->offset = base->obj_offset;
        F->G = F->opts.lanes / kL;
        F->lanes = F->opts.lanes;

        // The matrix in one segment (CSR of A, CSR of A^T = CSC of A), scaled
        // on the device by the single-LP engine's own set-up.
        std::vector<int> a_ptr, a_col, at_ptr(static_cast<size_t>(n) + 1);
        std::vector<double> a_val;
        csr_from_csc(base, a_ptr, a_col, a_val);
        for (int j = 0; j <= n; ++j) at_ptr[j] = static_cast<int>(base->col_start[j]);
        TileShape plain;
        upload_matrix(F->A, m, n, a_ptr, a_col.data(), a_val.data(), plain, F->s);
        upload_matrix(F->AT, n, m, at_ptr, base->row_index, base->value, plain, F->s);
        F->dR.alloc(m);
        F->dC.alloc(n);
        F->eta = pdlp_prepare_matrix(F->s, m, n, F->nnz, &F->A, &F->AT, F->dR.p, F->dC.p, F->R, F->C);
        {
            // The scaled column bounds are the same for every member.
            PdlpInstance I;
            pdlp_prepare_instance(base, F->prm, F->R, F->C, F->eta, F->mixed, I);
            F->lo.alloc(n);
            F->lo.upload(I.lo.data(), n, F->s);
            F->up.alloc(n);
            F->up.upload(I.up.data(), n, F->s);
        }

        const size_t L = static_cast<size_t>(F->lanes);
        const size_t NX = L * n, NY = L * m;
        for (auto* b : {&F->x, &F->xbar, &F->xa, &F->xp, &F->c, &F->qx, &F->rx}) b->alloc(NX);
        for (auto* b : {&F->y, &F->ya, &F->yp, &F->rl, &F->ru, &F->qy, &F->ry}) b->alloc(NY);
        if (F->mixed) {
            for (auto* b : {&F->fx, &F->fxbar, &F->fxp, &F->rc, &F->dl, &F->du}) b->alloc(NX);
            for (auto* b : {&F->fy, &F->fyp, &F->drl, &F->dru, &F->sa}) b->alloc(NY);
            F->GA.alloc(NX);
            F->HA.alloc(NY);
            F->aval32.alloc(static_cast<size_t>(F->nnz));
            F->tval32.alloc(static_cast<size_t>(F->nnz));
            const int gconv = static_cast<int>(std::max<long long>(1, std::min<long long>((F->nnz + 255) / 256, 4096)));
            k_fam_to_float<<<gconv, 256, 0, F->s>>>(F->A.val.p, F->aval32.p, F->nnz);
            k_fam_to_float<<<gconv, 256, 0, F->s>>>(F->AT.val.p, F->tval32.p, F->nnz);
            HIP_CHECK(hipGetLastError());
        }
        F->st.alloc(L);
        F->ltau.alloc(L);
        F->lsigma.alloc(L);
        F->lkb.alloc(L);
        F->ldone.alloc(L);
        F->lrestart.alloc(L);
        F->lfresh.alloc(L);
        {
            // Every lane starts empty (done, not fresh).
            std::vector<int> ones(L, 1), zeros(L, 0);
            std::vector<double> dz(L, 0.0);
            std::vector<PdlpState> sz(L);
            for (auto& z : sz) {
                z = PdlpState{};
                z.done = 1;
            }
            F->ldone.upload(ones.data(), L, F->s);
            F->lkb.upload(zeros.data(), L, F->s);
            F->lrestart.upload(zeros.data(), L, F->s);
            F->lfresh.upload(zeros.data(), L, F->s);
            F->ltau.upload(dz.data(), L, F->s);
            F->lsigma.upload(dz.data(), L, F->s);
            F->st.upload(sz.data(), L, F->s);
            HIP_CHECK(hipStreamSynchronize(F->s));
        }
        F->ebc = std::max(1, std::min(kEvalBlocks, (n + 255) / 256));
        F->ebr = std::max(1, std::min(kEvalBlocks, (m + 255) / 256));
        F->part_c.alloc(static_cast<size_t>(F->G) * F->ebc * 5 * kL);
        F->part_r.alloc(static_cast<size_t>(F->G) * F->ebr * 6 * kL);
        F->cert_c.alloc(static_cast<size_t>(F->G) * F->ebc * 6 * kL);
        F->cert_r.alloc(static_cast<size_t>(F->G) * F->ebr * 6 * kL);
        const size_t S = static_cast<size_t>(std::min(F->lanes, kStageMax));
        F->s_lanes.alloc(S);
        F->s_dst.alloc(S);
        F->s_c.alloc(S * n);
        F->s_x.alloc(S * n);
        F->s_rl.alloc(S * m);
        F->s_ru.alloc(S * m);
        F->s_y.alloc(S * m);
        F->s_st.alloc(S);
        F->s_src.alloc(2 * S);
        F->h_done.assign(L, 1);
        std::memset(&F->stats, 0, sizeof(F->stats));
        F->stats.lanes = F->lanes;
        F->stats.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = F.release();
        return 0;
    } catch (const std::exception& ex) {
        ctx->error = ex.what();
        if (F->s) (void)hipStreamSynchronize(F->s);
        return -1;
    }
}

NIR_API int nir_family_solve(nir_family* F, int64_t count, nir_family_fill_fn fill, nir_family_done_fn done,
                             void* user) {
    const nir::TraceRange trace("nir_family_solve");
    if (!F || !fill || !done || count < 0) return -1;
    try {
        HIP_CHECK(hipSetDevice(F->ctx->device));
        auto t0 = std::chrono::steady_clock::now();
        const int m = F->m, n = F->n;
        const size_t S = static_cast<size_t>(std::min(F->lanes, kStageMax));
        std::vector<long long> lane_inst(static_cast<size_t>(F->lanes), -1);
        // Free lanes, kept in descending order so that back() is the lowest:
        // refills fill the lowest groups first, and only the groups up to the
        // highest occupied lane are launched.
        std::vector<int> free_lanes;
        for (int l = F->lanes - 1; l >= 0; --l) free_lanes.push_back(l);
        auto update_groups = [&]() {
            int hi = -1;
            for (int l = F->lanes - 1; l >= 0; --l)
                if (lane_inst[l] >= 0) {
                    hi = l;
                    break;
                }
            F->Gl = hi < 0 ? 1 : hi / kL + 1;
        };
        update_groups();
        using clock = std::chrono::steady_clock;
        auto secs = [](clock::time_point a) { return std::chrono::duration<double>(clock::now() - a).count(); };
        long long next = 0, finished = 0;
        int active = 0;
        // Device start rows (nir_family_set_device_starts): member idx < rows
        // reads its start from row idx on the device.
        int dxn = 0, dyn = 0;
        const double *dx0 = nullptr, *dy0 = nullptr;
        {
            int len = 0, dev = 0;
            if (F->start_x) {
                dx0 = pod_rows(F->start_x, &dxn, &len, &dev);
                if (dx0 && (len != n || dev != F->ctx->device))
                    throw std::invalid_argument("device x starts: rows of the wrong length or device");
            }
            if (F->start_y) {
                dy0 = pod_rows(F->start_y, &dyn, &len, &dev);
                if (dy0 && (len != m || dev != F->ctx->device))
                    throw std::invalid_argument("device y starts: rows of the wrong length or device");
            }
            if (!dx0) dxn = 0;
            if (!dy0) dyn = 0;
        }
        std::vector<double> cost(n), rlo(m), rup(m), x0(n), y0(m);
        std::vector<Prepared> ready;
        // Prepare one instance on the host with the single-LP set-up code.
        auto prepare = [&](long long idx) {
            const auto tp = clock::now();
            int32_t warm = 0;
            double weight = 0.0;
            std::fill(x0.begin(), x0.end(), 0.0);
            std::fill(y0.begin(), y0.end(), 0.0);
            if (fill(user, idx, cost.data(), rlo.data(), rup.data(), x0.data(), y0.data(), &warm, &weight) != 0)
                throw std::runtime_error("family fill callback failed for instance " + std::to_string(idx));
            nir_lp lp{};
            lp.nrow = m;
            lp.ncol = n;
            lp.col_start = F->col_start.data();
            lp.row_index = F->row_index.data();
            lp.value = F->value.data();
            lp.cost = cost.data();
            lp.collo = F->collo.data();
            lp.colup = F->colup.data();
            lp.rowlo = rlo.data();
            lp.rowup = rup.data();
            lp.obj_offset = F->offset;
            nir_params p = F->prm;
            Prepared P;
            P.index = idx;
            P.dev_x = idx < dxn;
            P.dev_y = idx < dyn;
            p.x0 = ((warm & 1) && !P.dev_x) ? x0.data() : nullptr;
            p.y0 = ((warm & 2) && !P.dev_y) ? y0.data() : nullptr;
            if (weight > 0.0) p.primal_weight = weight;
            pdlp_prepare_instance(&lp, p, F->R, F->C, F->eta, F->mixed, P.inst);
            ready.push_back(std::move(P));
            F->stats.prepare_seconds += secs(tp);
        };
        std::vector<double> hc(S * n), hx(S * n), hrl(S * m), hru(S * m), hy(S * m);
        std::vector<PdlpState> hst(S);
        std::vector<int> hl(S);
        std::vector<long long> hsrc(2 * S);
        // Move prepared instances into free lanes (chunks of S), then
        // evaluate them at their start point.
        auto refill = [&]() {
            bool any = false;
            while (!ready.empty() && !free_lanes.empty()) {
                const size_t k = std::min({S, ready.size(), free_lanes.size()});
                for (size_t f = 0; f < k; ++f) {
                    Prepared& P = ready[ready.size() - k + f];
                    const int l = free_lanes.back();
                    free_lanes.pop_back();
                    hl[f] = l;
                    lane_inst[l] = P.index;
                    std::copy(P.inst.c.begin(), P.inst.c.end(), hc.begin() + f * n);
                    std::copy(P.inst.x0.begin(), P.inst.x0.end(), hx.begin() + f * n);
                    std::copy(P.inst.rl.begin(), P.inst.rl.end(), hrl.begin() + f * m);
                    std::copy(P.inst.ru.begin(), P.inst.ru.end(), hru.begin() + f * m);
                    std::copy(P.inst.y0.begin(), P.inst.y0.end(), hy.begin() + f * m);
                    hst[f] = P.inst.st;
                    hsrc[f] = P.dev_x ? P.index : -1;
                    hsrc[k + f] = P.dev_y ? P.index : -1;
                }
                ready.resize(ready.size() - k);
                F->s_lanes.upload(hl.data(), k, F->s);
                F->s_c.upload(hc.data(), k * n, F->s);
                F->s_x.upload(hx.data(), k * n, F->s);
                F->s_rl.upload(hrl.data(), k * m, F->s);
                F->s_ru.upload(hru.data(), k * m, F->s);
                F->s_y.upload(hy.data(), k * m, F->s);
                F->s_st.upload(hst.data(), k, F->s);
                F->s_src.upload(hsrc.data(), 2 * k, F->s);
                const int gx = std::max(1, std::min(64, (std::max(n, m) + 255) / 256));
                k_fam_scatter<<<dim3(gx, static_cast<unsigned>(k)), 256, 0, F->s>>>(
                    F->view(), static_cast<int>(k), F->s_lanes.p, F->s_c.p, F->s_x.p, F->s_rl.p, F->s_ru.p, F->s_y.p,
                    F->s_st.p, F->s_src.p, dx0, dy0);
                HIP_CHECK(hipGetLastError());
                // The staging buffers are reused by the next chunk.
                HIP_CHECK(hipStreamSynchronize(F->s));
                active += static_cast<int>(k);
                any = true;
            }
            if (any) {
                update_groups();
                if (F->mixed) launch_recenter<1>(F, F->view());
                launch_eval<1>(F, F->view());
                HIP_CHECK(hipGetLastError());
                ++F->stats.refills;
            }
        };
        // Lane l's member continued alone on the single-LP engine from the
        // lane's state (vectors and PdlpState copied on the device between
        // blocks, where both engines stand at the same point of the same
        // iteration), to its end; the result goes to the done callback and
        // the lane is emptied. Used for the hand-off of the last members and
        // for an FP32 phase that ends needing the FP64 tail.
        DevArray<double> ext;
        DevArray<float> ext32;
        std::vector<double> hcs, hxs, hys;
        bool single_stopped = false;
        auto run_single = [&](int l) {
            const size_t N = static_cast<size_t>(n), M = static_cast<size_t>(m);
            if (!ext.p) {
                ext.alloc(6 * N + 6 * M);
                if (F->mixed) ext32.alloc(6 * N + 5 * M);
                hcs.resize(N);
                hxs.resize(N);
                hys.resize(M);
            }
            const int gx = std::max(1, std::min(256, (std::max(n, m) + 255) / 256));
            if (F->mixed) k_fam_extract<1><<<gx, 256, 0, F->s>>>(F->view(), l, ext.p, ext32.p);
            else k_fam_extract<0><<<gx, 256, 0, F->s>>>(F->view(), l, ext.p, nullptr);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(hcs.data(), ext.p + 4 * N, sizeof(double) * N, hipMemcpyDeviceToHost, F->s));
            HIP_CHECK(hipStreamSynchronize(F->s));
            PdlpAdopt a;
            a.m = m;
            a.n = n;
            a.nnz = F->nnz;
            a.A = &F->A;
            a.AT = &F->AT;
            a.R = F->dR.p;
            a.C = F->dC.p;
            a.lo = F->lo.p;
            a.up = F->up.p;
            const double* ex = ext.p;
            const double* ey = ext.p + 5 * N;
            a.x = ex;
            a.xbar = ex + N;
            a.xa = ex + 2 * N;
            a.xp = ex + 3 * N;
            a.c = ex + 4 * N;
            a.y = ey;
            a.ya = ey + M;
            a.yp = ey + 2 * M;
            a.rl = ey + 3 * M;
            a.ru = ey + 4 * M;
            a.st = F->st.p + l;
            if (F->mixed) {
                a.mixed = true;
                a.GA = ext.p + 5 * N + 5 * M;
                a.HA = ext.p + 6 * N + 5 * M;
                const float* fxv = ext32.p;
                const float* fyv = ext32.p + 6 * N;
                a.fx = fxv;
                a.fxbar = fxv + N;
                a.fxp = fxv + 2 * N;
                a.rc = fxv + 3 * N;
                a.dl = fxv + 4 * N;
                a.du = fxv + 5 * N;
                a.fy = fyv;
                a.fyp = fyv + M;
                a.drl = fyv + 2 * M;
                a.dru = fyv + 3 * M;
                a.sa = fyv + 4 * M;
            }
            a.h_c.resize(N);
            for (int j = 0; j < n; ++j) a.h_c[j] = hcs[j] / F->C[j];   // caller units (C: powers of two)
            a.offset = F->offset;
            nir_info info;
            {
                std::unique_ptr<Pdlp, PdlpDeleter> P(pdlp_create());
                pdlp_adopt(*P, F->s, F->prm, a);
                pdlp_run(*P, F->s, F->prm, std::numeric_limits<long long>::max() / 4);
                pdlp_result(*P, F->s, hxs.data(), hys.data(), &info);
            }
            info.setup_seconds = F->stats.setup_seconds;
            info.solve_seconds = 0.0;
            if (info.status == 4) single_stopped = true;
            k_fam_release<<<1, 64, 0, F->s>>>(F->view(), l);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipStreamSynchronize(F->s));
            const long long idx = lane_inst[l];
            F->stats.total_iterations += info.iterations;
            ++finished;
            if (done(user, idx, hxs.data(), hys.data(), &info) != 0)
                throw std::runtime_error("family done callback failed for instance " + std::to_string(idx));
        };
        // Harvest every finished lane: T(z) unscaled, the objective in
        // double-double from the caller's costs (scaled cost / C, exact), the
        // status and counters as nir_session_result reports them.
        // Finished members fetched from the device and waiting for the host
        // to unscale them and hand them out (done while the next block runs).
        struct Fetched {
            std::vector<long long> index;
            std::vector<double> x, c, y;
            std::vector<PdlpState> st;
            std::vector<char> taken;   // already handed out (FP64 tail on the single-LP engine)
        };
        Fetched pending;
        long long fetched = 0;
        // Device half of a harvest: which lanes finished, and their T(z),
        // scaled costs and states copied to the host; the lanes are free
        // for refills at once.
        auto fetch = [&]() {
            const auto tw = clock::now();
            F->ldone.download(F->h_done.data(), F->h_done.size(), F->s);
            HIP_CHECK(hipStreamSynchronize(F->s));
            F->stats.wait_seconds += secs(tw);
            std::vector<int> fin;
            for (int l = 0; l < F->lanes; ++l)
                if (F->h_done[l] && lane_inst[l] >= 0) fin.push_back(l);
            Fetched out;
            out.x.resize(fin.size() * n);
            out.c.resize(fin.size() * n);
            out.y.resize(fin.size() * m);
            out.st.resize(fin.size());
            for (size_t b = 0; b < fin.size(); b += S) {
                const size_t k = std::min(S, fin.size() - b);
                std::copy(fin.begin() + b, fin.begin() + b + k, hl.begin());
                F->s_lanes.upload(hl.data(), k, F->s);
                const int gx = std::max(1, std::min(64, (std::max(n, m) + 255) / 256));
                if (F->mixed)
                    k_fam_gather<1><<<dim3(gx, static_cast<unsigned>(k)), 256, 0, F->s>>>(
                        F->view(), static_cast<int>(k), F->s_lanes.p, F->s_x.p, F->s_c.p, F->s_y.p, F->s_st.p);
                else
                    k_fam_gather<0><<<dim3(gx, static_cast<unsigned>(k)), 256, 0, F->s>>>(
                        F->view(), static_cast<int>(k), F->s_lanes.p, F->s_x.p, F->s_c.p, F->s_y.p, F->s_st.p);
                HIP_CHECK(hipGetLastError());
                F->s_x.download(out.x.data() + b * n, k * n, F->s);
                F->s_c.download(out.c.data() + b * n, k * n, F->s);
                F->s_y.download(out.y.data() + b * m, k * m, F->s);
                F->s_st.download(out.st.data() + b, k, F->s);
                HIP_CHECK(hipStreamSynchronize(F->s));
            }
            // An FP32 phase that met its tolerance with a polish to do, or
            // stalled, goes on in FP64: on the single-LP engine, from the
            // lane's state, before the lane is refilled.
            out.taken.assign(fin.size(), 0);
            for (size_t f = 0; f < fin.size(); ++f) {
                const PdlpState& h = out.st[f];
                if (F->mixed && ((h.status == 0 && F->prm.polish_tol > 0) || h.status == 2)) {
                    run_single(fin[f]);
                    out.taken[f] = 1;
                    ++F->stats.tails;
                }
            }
            for (int l : fin) {
                out.index.push_back(lane_inst[l]);
                lane_inst[l] = -1;
                free_lanes.push_back(l);
                --active;
            }
            fetched += static_cast<long long>(fin.size());
            if (!fin.empty()) {
                std::sort(free_lanes.begin(), free_lanes.end(), std::greater<int>());
                update_groups();
            }
            return out;
        };
        // Host half: T(z) unscaled, the objective in double-double from the
        // caller's costs (scaled cost / C, exact), the status and counters as
        // nir_session_result reports them, handed to the callback.
        auto process = [&](Fetched& B) {
            for (size_t f = 0; f < B.index.size(); ++f) {
                if (!B.taken.empty() && B.taken[f]) continue;
                double* xs = B.x.data() + f * n;
                double* ys = B.y.data() + f * m;
                const double* cs = B.c.data() + f * n;
                for (int j = 0; j < n; ++j) xs[j] *= F->C[j];
                for (int i = 0; i < m; ++i) ys[i] *= F->R[i];
                const PdlpState& h = B.st[f];
                nir_info info;
                std::memset(&info, 0, sizeof(info));
                info.status = sov_status(h);
                info.iterations = h.total;
                info.restarts = h.restarts;
                DDHost acc(F->offset);
                for (int j = 0; j < n; ++j) {
                    const double cj = cs[j] / F->C[j];
                    if (cj != 0.0 && xs[j] != 0.0) acc.add_prod(cj, xs[j]);
                }
                info.objective = acc.value();
                info.tau = h.tau;
                info.sigma = h.sigma;
                info.primal_weight = h.omega;
                info.rel_primal = h.rel_p;
                info.rel_dual = h.rel_d;
                info.rel_gap = h.rel_g;
                info.setup_seconds = F->stats.setup_seconds;
                F->stats.total_iterations += h.total;
                ++finished;
                if (done(user, B.index[f], xs, ys, &info) != 0)
                    throw std::runtime_error("family done callback failed for instance " +
                                             std::to_string(B.index[f]));
            }
            B = Fetched();
        };
        // Once no member waits for a lane, move the running members of the
        // highest groups into free lanes of the lowest ones, so that the
        // groups launched shrink with the members left.
        auto compact = [&]() {
            if (next < count || !ready.empty() || active == 0) return;
            const int need = (active + kL - 1) / kL;
            if (need >= F->Gl) return;
            std::vector<int> src, dst;
            for (int l = need * kL; l < F->lanes; ++l)
                if (lane_inst[l] >= 0) src.push_back(l);
            for (int l = 0; l < need * kL && dst.size() < src.size(); ++l)
                if (lane_inst[l] < 0) dst.push_back(l);
            if (src.empty() || dst.size() != src.size()) return;
            for (size_t b = 0; b < src.size(); b += S) {
                const size_t k = std::min(S, src.size() - b);
                std::copy(src.begin() + b, src.begin() + b + k, hl.begin());
                F->s_lanes.upload(hl.data(), k, F->s);
                std::vector<int> hd(dst.begin() + b, dst.begin() + b + k);
                F->s_dst.upload(hd.data(), k, F->s);
                const int gx = std::max(1, std::min(64, (std::max(n, m) + 255) / 256));
                if (F->mixed)
                    k_fam_move<1><<<dim3(gx, static_cast<unsigned>(k)), 256, 0, F->s>>>(
                        F->view(), static_cast<int>(k), F->s_lanes.p, F->s_dst.p);
                else
                    k_fam_move<0><<<dim3(gx, static_cast<unsigned>(k)), 256, 0, F->s>>>(
                        F->view(), static_cast<int>(k), F->s_lanes.p, F->s_dst.p);
                HIP_CHECK(hipGetLastError());
                HIP_CHECK(hipStreamSynchronize(F->s));
            }
            for (size_t f = 0; f < src.size(); ++f) {
                lane_inst[dst[f]] = lane_inst[src[f]];
                lane_inst[src[f]] = -1;
            }
            free_lanes.clear();
            for (int l = F->lanes - 1; l >= 0; --l)
                if (lane_inst[l] < 0) free_lanes.push_back(l);
            update_groups();
            ++F->stats.compactions;
        };
        bool stopped = false;
        // Once no member waits for a lane and at most opts.handoff still run,
        // each continues alone on the single-LP engine (run_single): a lone
        // member in a 64-lane group walks its rows one lane at a time, the
        // single-LP engine spreads each row's nonzeros over a workgroup.
        auto handoff = [&]() -> bool {
            process(pending);
            const long long before = fetched;
            const auto th = clock::now();
            for (int l = 0; l < F->lanes; ++l) {
                if (lane_inst[l] < 0) continue;
                // A stop request ends the remaining lanes in the family (the
                // main loop's stop path).
                if (F->prm.should_stop && F->prm.should_stop(F->prm.stop_user)) break;
                run_single(l);
                lane_inst[l] = -1;
                free_lanes.push_back(l);
                --active;
                ++fetched;
                ++F->stats.handoffs;
            }
            if (single_stopped) stopped = true;
            std::sort(free_lanes.begin(), free_lanes.end(), std::greater<int>());
            update_groups();
            F->stats.handoff_seconds += secs(th);
            return fetched > before;
        };
        // Every member ends by its iteration cap, so some lane must finish
        // within cap / eval_freq blocks of the last one that did.
        const long long drain_bound = std::max<long long>(1, F->prm.max_iters) / std::max(1, F->prm.eval_freq) + 8;
        long long blocks_idle = 0, fetched_seen = 0;
        while (fetched < count) {
            if (fetched != fetched_seen) {
                fetched_seen = fetched;
                blocks_idle = 0;
            } else if (blocks_idle > drain_bound) {
                throw std::runtime_error("family did not drain within its iteration bound");
            }
            // Keep enough prepared instances for every free lane.
            while (next < count && static_cast<long long>(ready.size()) < static_cast<long long>(free_lanes.size()))
                prepare(next++);
            refill();
            if (active == 0) {
                Fetched B = fetch();   // instances that finished at their start point
                process(B);
                continue;
            }
            if (F->prm.should_stop && F->prm.should_stop(F->prm.stop_user)) {
                k_fam_stop<<<(F->lanes + 255) / 256, 256, 0, F->s>>>(F->view(), F->lanes);
                HIP_CHECK(hipGetLastError());
                process(pending);
                Fetched B = fetch();
                process(B);
                stopped = true;
                break;
            }
            if (F->opts.handoff > 0 && next >= count && ready.empty() && active <= F->opts.handoff && handoff())
                continue;
            compact();
            run_block(F);
            ++F->stats.blocks;
            F->stats.group_blocks += F->Gl;
            ++blocks_idle;
            // While the block runs: hand out the previous block's members and
            // prepare the next ones.
            process(pending);
            while (next < count && static_cast<long long>(ready.size()) < std::max<long long>(1, F->lanes / 16))
                prepare(next++);
            pending = fetch();
        }
        process(pending);
        HIP_CHECK(hipStreamSynchronize(F->s));
        F->stats.instances += finished;
        F->stats.stopped = stopped ? 1 : 0;
        F->stats.solve_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    } catch (const std::exception& ex) {
        F->ctx->error = ex.what();
        (void)hipStreamSynchronize(F->s);
        return -1;
    }
}

NIR_API int nir_family_set_device_starts(nir_family* F, const nir_pod* px, const nir_pod* py) {
    if (!F) return -1;
    F->start_x = px;
    F->start_y = py;
    return 0;
}

NIR_API int nir_family_get_stats(nir_family* F, nir_family_stats* out) {
    if (!F || !out) return -1;
    *out = F->stats;
    return 0;
}

NIR_API void nir_family_destroy(nir_family* F) {
    if (!F) return;
    (void)hipSetDevice(F->ctx->device);
    (void)hipStreamSynchronize(F->s);
    delete F;
}

}  // extern "C"
