// Restarted PDHG ("pdlp" mode) on the device.
//
// Method (build-defined; restated bit for bit by the C oracle's or_pdlp):
//   * Diagonal preconditioning in the exponent domain: 10 Ruiz passes then one
//     Pock-Chambolle (alpha = 1) pass. Every factor is a power of two,
//     2^-floor((e+1)/2) for a row/column measure v in [2^e, 2^(e+1)), so
//     scaling and unscaling are exact and need no sqrt.
//   * Constant step: eta = 0.998 / sigma_max(A'), sigma_max from a power
//     iteration on A'^T A' (deterministic start vector, double-double norms).
//     tau = eta / omega, sigma = eta * omega for primal weight omega.
//   * Reflected Halpern iteration (Lu & Yang, "Restarted Halpern PDHG for
//     linear programming", 2024; reflection as in cuPDLPx, 2025): with T the
//     PDHG map of the reference loop (pdhg_solver.hpp:70-95) and k the
//     iterations since the last restart,
//         z_{k+1} = (k+1)/(k+2) (2 T(z_k) - z_k) + 1/(k+2) z_anchor.
//     The reflected primal point 2 x+ - x is exactly the extrapolated xbar the
//     dual step needs, so the anchor costs one extra read per element.
//   * Every eval_freq iterations (default 64, as in cuPDLP) the kernels keep
//     T(z) and evaluate in FP64 / double-double: primal residual of A x, dual
//     residual of c - A^T y, primal and dual objectives (PDLP's relative KKT,
//     in the caller's units), and the weighted KKT error
//         KKT_w = sqrt(w ||r_p||^2 + ||r_d||^2 / w + gap^2).
//     Restart (anchor and iterate := T(z)) when KKT_w <= 0.2 KKT_w(anchor), or
//     <= 0.8 KKT_w(anchor) while rising, or when the restart length reaches
//     0.36 of all iterations (cuPDLP's criteria); then
//     omega <- sqrt(omega * ||dy|| / ||dx||) (PDLP smoothing with theta 1/2).
//   * All decisions run in a one-workgroup kernel; the host replays a
//     hipGraph per block of eval_freq iterations and polls a done flag.
//   * Mixed precision: FP32 matrix values, iterates and accumulation in the
//     iteration kernels; FP64 matrix and arithmetic for every evaluation,
//     restart and step decision; then an FP64 tail from the FP32 point (restart
//     at it) down to the polish tolerance. An FP32 phase whose relative KKT
//     has not improved by 5% for max(100, evaluations/3) evaluations hands
//     over to FP64 early.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "devmat.hpp"
#include "pdlp.hpp"
#include "pdlp_core.hpp"

namespace nir {
namespace gpu {

#define PD_CHECK(expr)                                                                                \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------ reductions

// Reduce NQ dd accumulators over the workgroup (fixed tree), thread 0 writes.
template <int NQ>
__device__ __forceinline__ void block_dd_reduce(dd (&acc)[NQ], dd* sh, dd* out) {
    const int t = threadIdx.x;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) sh[q * kThreads + t] = acc[q];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int q = 0; q < NQ; ++q) sh[q * kThreads + t] = dd_add(sh[q * kThreads + t], sh[q * kThreads + t + w]);
        __syncthreads();
    }
    if (t == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) out[q] = sh[q * kThreads];
}

// ------------------------------------------------------------------ kernels

// The iteration kernels are compiled twice: LONG = 0 (the long-row code
// compiled out, for layouts without blocks of long or split rows - the
// short-row blocks then keep the registers and occupancy they have on their
// own) and LONG = 2 (every block, for layouts that have such blocks: they run
// beside the short-row blocks in the same launch).
template <int TILE, typename T, int LONG>
__global__ __launch_bounds__(kThreads) void k_primal_pdlp(Pass<T> P, const T* __restrict__ y, T* __restrict__ x,
                                                          T* __restrict__ xbar, const T* __restrict__ xa,
                                                          T* __restrict__ xp_out, const T* __restrict__ c,
                                                          const T* __restrict__ lo, const T* __restrict__ up,
                                                          const PdlpState* __restrict__ st, int k_lo, int save) {
    __shared__ T prod[lds_slots(TILE)];
    if (st->done) return;
    const T tau = static_cast<T>(st->tau);
    const Halpern<T> hw = halpern<T>(st->k_base + k_lo);
    struct Ops {
        T x, c, l, u, a;
    };
    // FP32 runs the re-centred form: the anchor is the FP64 point the deltas
    // are taken from, its delta is 0 and no anchor is read.
    constexpr bool kDelta = std::is_same<T, float>::value;
    const auto pre = [&](int j) {
        Ops o{T(0), T(0), T(0), T(0), T(0)};
        if (P.last) o = Ops{x[j], c[j], lo[j], up[j], kDelta ? T(0) : xa[j]};
        return o;
    };
    stream_rows<TILE, 1, LONG, LONG == 2 ? 16 : 8>(P, y, prod, blockIdx.x, pre, [&](int j, T g, const Ops& o) {
        const PrimalOut<T> r = primal_point<T>(o.x, g, o.c, o.l, o.u, o.a, tau, hw);
        xbar[j] = r.xb;
        x[j] = r.next;
        if (save) xp_out[j] = r.cand;
    });
}

template <int TILE, typename T, int LONG>
__global__ __launch_bounds__(kThreads) void k_dual_pdlp(Pass<T> P, const T* __restrict__ xbar, T* __restrict__ y,
                                                        const T* __restrict__ ya, T* __restrict__ yp_out,
                                                        const T* __restrict__ rl, const T* __restrict__ ru,
                                                        const PdlpState* __restrict__ st, int k_lo, int save) {
    __shared__ T prod[lds_slots(TILE)];
    if (st->done) return;
    const T sigma = static_cast<T>(st->sigma);
    const Halpern<T> hw = halpern<T>(st->k_base + k_lo);
    struct Ops {
        T y, l, u, a;
    };
    const auto pre = [&](int i) {
        Ops o{T(0), T(0), T(0), T(0)};
        if (P.last) o = Ops{y[i], rl[i], ru[i], ya[i]};
        return o;
    };
    // FP32 runs the re-centred form; `ya` then holds sa = -Y / sigma.
    constexpr bool kDelta = std::is_same<T, float>::value;
    stream_rows<TILE, 1, LONG, LONG == 2 ? 16 : 8>(P, xbar, prod, blockIdx.x, pre, [&](int i, T h, const Ops& o) {
        const DualOut<T> r = kDelta ? dual_point_delta<T>(o.y, h, o.l, o.u, o.a, sigma, hw)
                                    : dual_point<T>(o.y, h, o.l, o.u, o.a, sigma, hw);
        y[i] = r.next;
        if (save) yp_out[i] = r.yp;
    });
}

// Column evaluation: lambda = c - A^T yp, its residual against the sign the
// column bounds allow, the column part of the dual objective, c^T xp and
// ||xp - x_anchor||^2. Quantities: 0 rd^2, 1 (rd/C)^2, 2 c^T x, 3 dual-obj
// column terms, 4 ||dx||^2.
template <int TILE, typename TI>
__global__ __launch_bounds__(kThreads) void k_eval_cols(Pass<double> P, const TI* __restrict__ yp,
                                                        const TI* __restrict__ xp, const TI* __restrict__ xa,
                                                        const double* __restrict__ c, const double* __restrict__ lo,
                                                        const double* __restrict__ up,
                                                        const double* __restrict__ cscale,
                                                        const double* __restrict__ anchor,
                                                        const double* __restrict__ base,
                                                        const PdlpState* __restrict__ st, dd* __restrict__ part,
                                                        int force) {
    __shared__ double prod[lds_slots(TILE)];
    __shared__ dd sh[5 * kThreads];
    if (st->done && !force) return;
    dd acc[5] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    // Re-centred iterates (anchor != null): T(z) = anchor + delta and
    // A^T yp = base + A^T dyp.
    stream_rows<TILE, 1>(P, yp, prod, blockIdx.x, [&](int j, double g) {
        if (anchor)
            col_terms(base[j] + g, c[j], lo[j], up[j], cscale[j], anchor[j] + static_cast<double>(xp[j]), anchor[j],
                      acc);
        else
            col_terms(g, c[j], lo[j], up[j], cscale[j], static_cast<double>(xp[j]), static_cast<double>(xa[j]), acc);
    });
    if (P.last) block_dd_reduce<5>(acc, sh, part + 5 * blockIdx.x);
}

// Row evaluation: r_p = A xp - clamp(A xp), sign violations of yp, the row
// part of the dual objective and ||yp - y_anchor||^2. Quantities: 0 rp^2,
// 1 (rp/R)^2, 2 yviol^2, 3 (R yviol)^2, 4 dual-obj row terms, 5 ||dy||^2.
template <int TILE, typename TI>
__global__ __launch_bounds__(kThreads) void k_eval_rows(Pass<double> P, const TI* __restrict__ xp,
                                                        const TI* __restrict__ yp, const TI* __restrict__ ya,
                                                        const double* __restrict__ rl, const double* __restrict__ ru,
                                                        const double* __restrict__ rscale,
                                                        const double* __restrict__ anchor,
                                                        const double* __restrict__ base,
                                                        const PdlpState* __restrict__ st, dd* __restrict__ part,
                                                        int force) {
    __shared__ double prod[lds_slots(TILE)];
    __shared__ dd sh[6 * kThreads];
    if (st->done && !force) return;
    dd acc[6] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    stream_rows<TILE, 1>(P, xp, prod, blockIdx.x, [&](int i, double h) {
        if (anchor)
            row_terms(base[i] + h, rl[i], ru[i], rscale[i], anchor[i] + static_cast<double>(yp[i]), anchor[i], acc);
        else
            row_terms(h, rl[i], ru[i], rscale[i], static_cast<double>(yp[i]), static_cast<double>(ya[i]), acc);
    });
    if (P.last) block_dd_reduce<6>(acc, sh, part + 6 * blockIdx.x);
}

// Decision kernel (one workgroup): reduce the partials, evaluate KKT,
// terminate / restart / update the primal weight. mode 0: regular
// evaluation after a block of eval_freq iterations; mode 1: initial
// evaluation of the start point (sets the anchor KKT, may terminate).
__global__ __launch_bounds__(kThreads) void k_decide(PdlpState* st, const dd* __restrict__ pc, int nbc,
                                                     const dd* __restrict__ pr, int nbr, int mode) {
    __shared__ dd sh[11 * kThreads];
    if (st->done && mode == 0) return;
    dd acc[11];
    for (int q = 0; q < 11; ++q) acc[q] = dd_zero();
    for (int b = threadIdx.x; b < nbc; b += kThreads)
        for (int q = 0; q < 5; ++q) acc[q] = dd_add(acc[q], pc[5 * b + q]);
    for (int b = threadIdx.x; b < nbr; b += kThreads)
        for (int q = 0; q < 6; ++q) acc[5 + q] = dd_add(acc[5 + q], pr[6 * b + q]);
    __shared__ dd tot[11];
    block_dd_reduce<11>(acc, sh, tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    pdlp_decide(*st, tot, mode);
}

// Restart: iterate and anchor become T(z).
template <typename T>
__global__ void k_restart_copy(const PdlpState* __restrict__ st, int n, int m, const T* __restrict__ xp,
                               const T* __restrict__ yp, T* __restrict__ x, T* __restrict__ xa, T* __restrict__ y,
                               T* __restrict__ ya, int force) {
    if (!force && (st->done || !st->restart)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int j = t; j < n; j += stride) {
        const T v = xp[j];
        x[j] = v;
        xa[j] = v;
    }
    for (int i = t; i < m; i += stride) {
        const T v = yp[i];
        y[i] = v;
        ya[i] = v;
    }
}

// ------------------------------------------------------------------ infeasibility certificates

// The step between the anchors of a restart is tested as a ray (the terms and
// the decision are in pdlp_core.hpp). Only the kernels' early exits run when
// the evaluation did not restart.
__device__ __forceinline__ bool cert_active(const PdlpState* st) { return !st->done && st->restart; }

// Before the anchor moves: keep it.
__global__ void k_cert_save(const PdlpState* __restrict__ st, int n, int m, const double* __restrict__ X,
                            const double* __restrict__ Y, double* __restrict__ qx, double* __restrict__ qy) {
    if (!cert_active(st)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (int j = t; j < n; j += stride) qx[j] = X[j];
    for (int i = t; i < m; i += stride) qy[i] = Y[i];
}

// After it moved: d = new anchor - old anchor.
__global__ void k_cert_diff(const PdlpState* __restrict__ st, int n, int m, const double* __restrict__ X,
                            const double* __restrict__ Y, const double* __restrict__ qx,
                            const double* __restrict__ qy, double* __restrict__ dx, double* __restrict__ dy) {
    if (!cert_active(st)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (int j = t; j < n; j += stride) dx[j] = X[j] - qx[j];
    for (int i = t; i < m; i += stride) dy[i] = Y[i] - qy[i];
}

// Lines (rows of A or of A^T) longer than `long_len` nonzeros are split
// rows of a one-segment layout; the one-lane kernels skip them and the
// split-row kernels below cover them.
__device__ __forceinline__ bool split_line(int nseg, const int* __restrict__ rp, int r, int long_len) {
    return nseg == 1 && rp[r + 1] - rp[r] > long_len;
}

// Column terms (one lane per column, (A^T dy)_j over CSR(A^T)).
__global__ __launch_bounds__(kThreads) void k_cert_cols(const PdlpState* __restrict__ st, int n, int nseg,
                                                        const int* __restrict__ rp, const int* __restrict__ col,
                                                        const double* __restrict__ val,
                                                        const double* __restrict__ dy, const double* __restrict__ dx,
                                                        const double* __restrict__ c, const double* __restrict__ lo,
                                                        const double* __restrict__ up, dd* __restrict__ part,
                                                        int long_len) {
    __shared__ dd sh[6 * kThreads];
    if (!cert_active(st)) return;
    dd acc[6] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    const int j = blockIdx.x * kThreads + threadIdx.x;
    if (j < n && !split_line(nseg, rp, j, long_len))
        cert_col_terms(-dd_val(row_dot_dd(n, nseg, rp, col, val, dy, j)), c[j], lo[j], up[j], dx[j], acc);
    block_dd_reduce<6>(acc, sh, part + 6 * blockIdx.x);
}

// Row terms (one lane per row, (A dx)_i over CSR(A)).
__global__ __launch_bounds__(kThreads) void k_cert_rows(const PdlpState* __restrict__ st, int m, int nseg,
                                                        const int* __restrict__ rp, const int* __restrict__ col,
                                                        const double* __restrict__ val,
                                                        const double* __restrict__ dx, const double* __restrict__ dy,
                                                        const double* __restrict__ rl, const double* __restrict__ ru,
                                                        dd* __restrict__ part, int long_len) {
    __shared__ dd sh[6 * kThreads];
    if (!cert_active(st)) return;
    dd acc[6] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < m && !split_line(nseg, rp, i, long_len))
        cert_row_terms(dd_val(row_dot_dd(m, nseg, rp, col, val, dx, i)), rl[i], ru[i], dy[i], acc);
    block_dd_reduce<6>(acc, sh, part + 6 * blockIdx.x);
}

// Double-double products of the split rows, one workgroup per piece (other
// blocks exit): chunks of kChunk products summed in dd from zero by one lane
// each, then each group's chunk sums in dd by lane 0 into its slot of `gdd`.
// A piece holds at most 2 * kGroupLen products (64 chunks).
__global__ __launch_bounds__(kThreads) void k_cert_pieces(const PdlpState* __restrict__ st, Csr<double> M,
                                                          const double* __restrict__ v, dd* __restrict__ gdd) {
    __shared__ dd cs[2 * kGroupChunks];
    if (!cert_active(st)) return;
    const int4 d = M.desc[blockIdx.x];
    if (d.y >= 0) return;
    const int4 L = M.lrow[-d.y - 1];
    const int tid = threadIdx.x;
    const int len = d.w - d.z, nch = (len + kChunk - 1) / kChunk;
    if (tid < nch) {
        dd a = dd_zero();
        const int q1 = min((tid + 1) * kChunk, len);
        for (int q = d.z + tid * kChunk; q < d.z + q1; ++q) a = dd_add(a, dd_prod(M.val[q], v[M.col[q]]));
        cs[tid] = a;
    }
    __syncthreads();
    if (tid == 0) {
        int g = L.x + (d.z - M.rowptr[L.w]) / kGroupLen;
        for (int c0 = 0; c0 < nch; c0 += kGroupChunks, ++g) {
            dd a = dd_zero();
            for (int c = c0; c < min(nch, c0 + kGroupChunks); ++c) a = dd_add(a, cs[c]);
            gdd[g] = a;
        }
    }
}

// A split row's dd product: its group sums added in dd from zero in order.
__device__ __forceinline__ double split_dot(const int4& L, const dd* __restrict__ gdd) {
    dd a = dd_zero();
    for (int g = 0; g < L.y; ++g) a = dd_add(a, gdd[L.x + g]);
    return dd_val(a);
}

// Row terms of the split rows of A (one lane per split row).
__global__ __launch_bounds__(kThreads) void k_cert_split_rows(const PdlpState* __restrict__ st,
                                                              const int4* __restrict__ lrow, int nsplit,
                                                              const dd* __restrict__ gdd,
                                                              const double* __restrict__ dy,
                                                              const double* __restrict__ rl,
                                                              const double* __restrict__ ru, dd* __restrict__ part) {
    __shared__ dd sh[6 * kThreads];
    if (!cert_active(st)) return;
    dd acc[6] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k < nsplit) {
        const int4 L = lrow[k];
        const int i = L.w;
        cert_row_terms(split_dot(L, gdd), rl[i], ru[i], dy[i], acc);
    }
    block_dd_reduce<6>(acc, sh, part + 6 * blockIdx.x);
}

// Column terms of the split rows of A^T (one lane per split column).
__global__ __launch_bounds__(kThreads) void k_cert_split_cols(const PdlpState* __restrict__ st,
                                                              const int4* __restrict__ lrow, int nsplit,
                                                              const dd* __restrict__ gdd,
                                                              const double* __restrict__ dx,
                                                              const double* __restrict__ c,
                                                              const double* __restrict__ lo,
                                                              const double* __restrict__ up, dd* __restrict__ part) {
    __shared__ dd sh[6 * kThreads];
    if (!cert_active(st)) return;
    dd acc[6] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k < nsplit) {
        const int4 L = lrow[k];
        const int j = L.w;
        cert_col_terms(-split_dot(L, gdd), c[j], lo[j], up[j], dx[j], acc);
    }
    block_dd_reduce<6>(acc, sh, part + 6 * blockIdx.x);
}

// Reduce both partial sets and decide (one workgroup).
__global__ __launch_bounds__(kThreads) void k_cert_decide(PdlpState* st, const dd* __restrict__ pc, int nbc,
                                                          const dd* __restrict__ pr, int nbr) {
    __shared__ dd sh[6 * kThreads];
    __shared__ dd tot[6];
    if (!cert_active(st)) return;
    dd acc[6] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    for (int b = threadIdx.x; b < nbc; b += kThreads)
        for (int q = 0; q < 6; ++q) acc[q] = dd_add(acc[q], pc[6 * b + q]);
    for (int b = threadIdx.x; b < nbr; b += kThreads)
        for (int q = 0; q < 6; ++q) acc[q] = dd_add(acc[q], pr[6 * b + q]);
    block_dd_reduce<6>(acc, sh, tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int code = cert_decide(tot);
    if (code) {
        st->done = 1;
        st->status = code;
        st->restart = 0;
    }
}

// Plain y = M v (power iteration).
template <int TILE>
__global__ __launch_bounds__(kThreads) void k_spmv_plain(Pass<double> P, const double* __restrict__ v,
                                                         double* __restrict__ out) {
    __shared__ double prod[lds_slots(TILE)];
    stream_rows<TILE, 1>(P, v, prod, blockIdx.x, [&](int r, double s) { out[r] = s; });
}

// y = M v only when the decision restarted (or `force`): the FP64 anchor
// products of the re-centred FP32 phase.
template <int TILE>
__global__ __launch_bounds__(kThreads) void k_spmv_on_restart(Pass<double> P, const double* __restrict__ v,
                                                              double* __restrict__ out,
                                                              const PdlpState* __restrict__ st, int force) {
    __shared__ double prod[lds_slots(TILE)];
    if (!force && (st->done || !st->restart)) return;
    stream_rows<TILE, 1>(P, v, prod, blockIdx.x, [&](int r, double s) { out[r] = s; });
}

// Re-centring, step 1: the anchor moves to T(z) = anchor + delta (FP64).
__global__ void k_anchor_move(const PdlpState* __restrict__ st, int n, int m, double* X, double* Y,
                              const float* __restrict__ dxp, const float* __restrict__ dyp, int force) {
    if (!force && (st->done || !st->restart)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (int j = t; j < n; j += stride) X[j] = X[j] + static_cast<double>(dxp[j]);
    for (int i = t; i < m; i += stride) Y[i] = Y[i] + static_cast<double>(dyp[i]);
}

// Re-centring, step 3 (after HA = A X, GA = A^T Y): FP32 constants re-based
// on the anchor and zero deltas.
__global__ void k_recenter(const PdlpState* __restrict__ st, int n, int m, const double* __restrict__ c,
                           const double* __restrict__ lo, const double* __restrict__ up,
                           const double* __restrict__ rl, const double* __restrict__ ru,
                           const double* __restrict__ X, const double* __restrict__ Y,
                           const double* __restrict__ HA, const double* __restrict__ GA, float* rc, float* dl,
                           float* du, float* drl, float* dru, float* sa, float* dx, float* dxp, float* dy, float* dyp,
                           int force) {
    if (!force && (st->done || !st->restart)) return;
    const double sigma = st->sigma;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (int j = t; j < n; j += stride) {
        rc[j] = static_cast<float>(c[j] - GA[j]);
        dl[j] = static_cast<float>(lo[j] - X[j]);
        du[j] = static_cast<float>(up[j] - X[j]);
        dx[j] = 0.0f;
        dxp[j] = 0.0f;
    }
    for (int i = t; i < m; i += stride) {
        drl[i] = static_cast<float>(rl[i] - HA[i]);
        dru[i] = static_cast<float>(ru[i] - HA[i]);
        sa[i] = static_cast<float>(-Y[i] / sigma);
        dy[i] = 0.0f;
        dyp[i] = 0.0f;
    }
}

// out = anchor + delta (FP64), the FP32 phase's point in FP64.
__global__ void k_delta_to_double(const double* anchor, const float* delta, double* out, long long n) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = t; i < n; i += stride) out[i] = anchor[i] + static_cast<double>(delta[i]);
}

// Row measures over a segmented CSR matrix, segments in column order: max |a|
// (what = 0) or sum |a| (what = 1). Split rows (more than `long_len`
// nonzeros) get 0 here; k_piece_measure and k_split_total finish them.
__global__ void k_row_measure(const int* __restrict__ rowptr, int nseg, int nrows, const double* __restrict__ val,
                              double* out, int what, int long_len) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double acc = 0.0;
    if (split_line(nseg, rowptr, r, long_len)) {
        out[r] = 0.0;
        return;
    }
    for (int s = 0; s < nseg; ++s) {
        const int* rp = rowptr + static_cast<size_t>(s) * (nrows + 1);
        for (int p = rp[r]; p < rp[r + 1]; ++p) {
            const double a = fabs(val[p]);
            acc = what == 0 ? amax(acc, a) : acc + a;
        }
    }
    out[r] = acc;
}

__global__ void k_factor(const double* meas, double* f, double* total, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double g = pow2_inv_sqrt(meas[r]);
    f[r] = g;
    total[r] *= g;
}

// Split rows, one workgroup per piece (other blocks exit). what = 0: the
// piece's max |a| joins out[row] by an atomic max on the bit patterns
// (nonnegative doubles order as their bits; out[row] starts at 0). what = 1:
// sum |a| in the chunked order of the products (chunks of kChunk from 0.0 by
// one lane each, each group's chunk sums from 0.0 by lane 0) into the row's
// group slots of M.gsum; k_split_total adds the groups.
__global__ __launch_bounds__(kThreads) void k_piece_measure(Csr<double> M, double* out, int what) {
    __shared__ double red[2 * kGroupChunks];
    const int4 d = M.desc[blockIdx.x];
    if (d.y >= 0) return;
    const int4 L = M.lrow[-d.y - 1];
    const int tid = threadIdx.x;
    if (what == 0) {
        double a = 0.0;
        for (int p = d.z + tid; p < d.w; p += kThreads) a = amax(a, fabs(M.val[p]));
        a = block_max(a, red);
        if (tid == 0)
            atomicMax(reinterpret_cast<unsigned long long*>(out + L.w),
                      static_cast<unsigned long long>(__double_as_longlong(a)));
        return;
    }
    const int len = d.w - d.z, nch = (len + kChunk - 1) / kChunk;
    if (tid < nch) {
        double a = 0.0;
        const int q1 = min((tid + 1) * kChunk, len);
        for (int q = d.z + tid * kChunk; q < d.z + q1; ++q) a += fabs(M.val[q]);
        red[tid] = a;
    }
    __syncthreads();
    if (tid == 0) {
        int g = L.x + (d.z - M.rowptr[L.w]) / kGroupLen;
        for (int c0 = 0; c0 < nch; c0 += kGroupChunks, ++g) {
            double a = 0.0;
            for (int c = c0; c < min(nch, c0 + kGroupChunks); ++c) a += red[c];
            M.gsum[g] = a;
        }
    }
}

// out[row] of each split row = its group sums added from 0.0 in order.
__global__ void k_split_total(const int4* __restrict__ lrow, int nsplit, const double* __restrict__ gsum,
                              double* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsplit) return;
    const int4 L = lrow[k];
    double a = 0.0;
    for (int g = 0; g < L.y; ++g) a += gsum[L.x + g];
    out[L.w] = a;
}

// The pieces' share of k_scale_csr (one workgroup per piece).
__global__ __launch_bounds__(kThreads) void k_piece_scale(Csr<double> M, double* val, const double* rf,
                                                          const double* cf) {
    const int4 d = M.desc[blockIdx.x];
    if (d.y >= 0) return;
    const double a = rf[M.lrow[-d.y - 1].w];
    for (int p = d.z + threadIdx.x; p < d.w; p += kThreads) val[p] = val[p] * (a * cf[M.col[p]]);
}

// val[p] *= rf[row] * cf[col[p]] over every segment (all factors powers of
// two); split rows (more than `long_len` nonzeros) are left to k_piece_scale.
__global__ void k_scale_csr(const int* __restrict__ rowptr, int nseg, int nrows, const int* __restrict__ col,
                            double* val, const double* rf, const double* cf, int long_len) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows || split_line(nseg, rowptr, r, long_len)) return;
    const double a = rf[r];
    for (int s = 0; s < nseg; ++s) {
        const int* rp = rowptr + static_cast<size_t>(s) * (nrows + 1);
        for (int p = rp[r]; p < rp[r + 1]; ++p) val[p] = val[p] * (a * cf[col[p]]);
    }
}

__global__ void k_to_float(const double* in, float* out, long long n) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = t; i < n; i += stride) out[i] = static_cast<float>(in[i]);
}

__global__ void k_to_double(const float* in, double* out, long long n) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = t; i < n; i += stride) out[i] = static_cast<double>(in[i]);
}

// Sum of squares (dd) per block for the power iteration.
__global__ __launch_bounds__(kThreads) void k_sumsq(const double* v, int n, dd* part) {
    __shared__ dd sh[kThreads];
    dd acc[1] = {dd_zero()};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
        acc[0] = dd_add(acc[0], dd_prod(v[i], v[i]));
    block_dd_reduce<1>(acc, sh, part + blockIdx.x);
}

__global__ void k_scale_vec(double* v, int n, double s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = v[i] / s;
}

// ------------------------------------------------------------------ host side

template <typename T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        if (p) (void)hipFree(p);
        n = count;
        PD_CHECK(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T)));
    }
    void up(const T* src, size_t count, hipStream_t s) {
        if (count) PD_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void down(T* dst, size_t count, hipStream_t s) const {
        if (count) PD_CHECK(hipMemcpyAsync(dst, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    void from_device(const T* src, size_t count, hipStream_t s) {
        if (count) PD_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
};

template <typename T>
struct Iterates {
    Buf<T> x, xbar, xa, xp, y, ya, yp;
    Buf<T> c, lo, up, rl, ru;
    void alloc(int n, int m) {
        x.alloc(n); xbar.alloc(n); xa.alloc(n); xp.alloc(n);
        y.alloc(m); ya.alloc(m); yp.alloc(m);
        c.alloc(n); lo.alloc(n); up.alloc(n); rl.alloc(m); ru.alloc(m);
    }
};

// FP32 copy of a matrix's values (same segmented structure as the FP64 one).
struct Values32 {
    Buf<float> val, part;
};

struct Pdlp {
    int n = 0, m = 0;
    long long nnz = 0;
    DevMatrix* A = nullptr;    // FP64, scaled in place (owned by the session)
    DevMatrix* AT = nullptr;
    Values32 a32, at32;        // FP32 copies of the scaled values (mixed mode)
    Buf<double> R, C;          // engine scaling (powers of two)
    Buf<double> Rtot, Ctot;    // engine x caller scaling, for KKT in the caller's units
    Iterates<double> d;        // FP32 phase: d.xa / d.ya are the FP64 anchor X, Y
    Iterates<float> f;         // FP32 phase: deltas and re-based constants (f.ya = -Y / sigma)
    Buf<double> HA, GA;        // FP32 phase: A X and A^T Y of the anchor
    Buf<double> qx, qy;        // certificate: the anchor before the restart
    Buf<double> rx, ry;        // certificate: the step between the two anchors
    Buf<dd> cert_c, cert_r;    // certificate: per-block partial sums
    Buf<dd> cert_gt, cert_ga;  // certificate: group sums of the split rows of A^T and A
    Buf<dd> part_c, part_r;
    Buf<PdlpState> st;
    PdlpState* h_st = nullptr;   // pinned, two poll slots
    hipEvent_t ev[2] = {nullptr, nullptr};
    int pending[2] = {0, 0};
    long long serial = 0;
    bool mixed = false;          // current phase runs FP32 kernels
    bool polish_done = false;
    double polish_tol = 0.0, tol = 1e-6;
    int eval_freq = 64;
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool graph_mixed = false;
    bool finished = false;
    int status = 0;
    long long fp64_iters_at_switch = -1;
    double setup_s = 0.0, solve_s = 0.0;
    std::vector<double> h_c;
    double offset = 0.0;
    int (*should_stop)(void*) = nullptr;
    void* stop_user = nullptr;
    long long polls = 0;

    ~Pdlp() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (h_st) (void)hipHostFree(h_st);
    }
};

Pdlp* pdlp_create() { return new Pdlp; }
void PdlpDeleter::operator()(Pdlp* p) const { delete p; }

namespace {

int grid_for(long long n) { return static_cast<int>(std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192))); }

// Pass `s` of a matrix product in precision T.
template <typename T>
Pass<T> pass_of(const DevMatrix& M, const Values32& v32, int s);
template <>
Pass<double> pass_of<double>(const DevMatrix& M, const Values32&, int s) {
    return M.shape.pass(M.val.p, M.part.p, s);
}
template <>
Pass<float> pass_of<float>(const DevMatrix& M, const Values32& v32, int s) {
    return M.shape.pass(static_cast<const float*>(v32.val.p), v32.part.p, s);
}

template <typename T>
void launch_primal(Pdlp& P, hipStream_t s, Iterates<T>& it, int k_lo, int save) {
    const DevMatrix& M = *P.AT;
    for (int q = 0; q < M.shape.nseg; ++q) {
        const Pass<T> ps = pass_of<T>(M, P.at32, q);
        // A second launch for the long-row blocks alone would follow the
        // short-row launch instead of overlapping it: measured slower (a
        // 4,000-nonzero row: +20 us per iteration; a split row of 600,000:
        // 67 vs 56 us, a forked stream under graph capture no better), so
        // such layouts take one LONG = 2 launch.
        if (M.shape.nlong[q] > 0) {
            TILE_DISPATCH(M.shape.shape.tile,
                          (k_primal_pdlp<TL, T, 2><<<ps.M.nblk, kThreads, 0, s>>>(
                              ps, it.y.p, it.x.p, it.xbar.p, it.xa.p, it.xp.p, it.c.p, it.lo.p, it.up.p, P.st.p, k_lo,
                              save)));
        } else {
            TILE_DISPATCH(M.shape.shape.tile,
                          (k_primal_pdlp<TL, T, 0><<<ps.M.nblk, kThreads, 0, s>>>(
                              ps, it.y.p, it.x.p, it.xbar.p, it.xa.p, it.xp.p, it.c.p, it.lo.p, it.up.p, P.st.p, k_lo,
                              save)));
        }
    }
}

template <typename T>
void launch_dual(Pdlp& P, hipStream_t s, Iterates<T>& it, int k_lo, int save) {
    const DevMatrix& M = *P.A;
    for (int q = 0; q < M.shape.nseg; ++q) {
        const Pass<T> ps = pass_of<T>(M, P.a32, q);
        if (M.shape.nlong[q] > 0) {
            TILE_DISPATCH(M.shape.shape.tile,
                          (k_dual_pdlp<TL, T, 2><<<ps.M.nblk, kThreads, 0, s>>>(
                              ps, it.xbar.p, it.y.p, it.ya.p, it.yp.p, it.rl.p, it.ru.p, P.st.p, k_lo, save)));
        } else {
            TILE_DISPATCH(M.shape.shape.tile,
                          (k_dual_pdlp<TL, T, 0><<<ps.M.nblk, kThreads, 0, s>>>(
                              ps, it.xbar.p, it.y.p, it.ya.p, it.yp.p, it.rl.p, it.ru.p, P.st.p, k_lo, save)));
        }
    }
}

template <typename T>
void launch_iter(Pdlp& P, hipStream_t s, Iterates<T>& it, int k_lo, int save) {
    launch_primal<T>(P, s, it, k_lo, save);
    launch_dual<T>(P, s, it, k_lo, save);
    PD_CHECK(hipGetLastError());
}

// Re-centre the FP32 phase on its anchor: move the anchor to T(z) (unless
// `force`, used at setup where the deltas are zero), recompute A X and A^T Y
// in FP64 and re-base the FP32 constants. Every kernel returns at once unless
// the decision restarted, so the sequence can sit in every captured block.
void recenter(Pdlp& P, hipStream_t s, int force) {
    const int n = P.n, m = P.m;
    if (!force)
        k_anchor_move<<<grid_for(std::max(n, m)), 256, 0, s>>>(P.st.p, n, m, P.d.xa.p, P.d.ya.p, P.f.xp.p, P.f.yp.p,
                                                               0);
    for (int q = 0; q < P.A->shape.nseg; ++q) {
        const Pass<double> ps = P.A->shape.pass(P.A->val.p, P.A->part.p, q);
        TILE_DISPATCH(P.A->shape.shape.tile,
                      k_spmv_on_restart<TL><<<ps.M.nblk, kThreads, 0, s>>>(ps, P.d.xa.p, P.HA.p, P.st.p, force));
    }
    for (int q = 0; q < P.AT->shape.nseg; ++q) {
        const Pass<double> ps = P.AT->shape.pass(P.AT->val.p, P.AT->part.p, q);
        TILE_DISPATCH(P.AT->shape.shape.tile,
                      k_spmv_on_restart<TL><<<ps.M.nblk, kThreads, 0, s>>>(ps, P.d.ya.p, P.GA.p, P.st.p, force));
    }
    k_recenter<<<grid_for(std::max(n, m)), 256, 0, s>>>(P.st.p, n, m, P.d.c.p, P.d.lo.p, P.d.up.p, P.d.rl.p, P.d.ru.p,
                                                        P.d.xa.p, P.d.ya.p, P.HA.p, P.GA.p, P.f.c.p, P.f.lo.p,
                                                        P.f.up.p, P.f.rl.p, P.f.ru.p, P.f.ya.p, P.f.x.p, P.f.xp.p,
                                                        P.f.y.p, P.f.yp.p, force);
    PD_CHECK(hipGetLastError());
}

// Certificate test of the step between the anchors (kernels exit early unless
// the decision just restarted).
void launch_certificate(Pdlp& P, hipStream_t s) {
    const int n = P.n, m = P.m;
    k_cert_diff<<<grid_for(std::max(n, m)), 256, 0, s>>>(P.st.p, n, m, P.d.xa.p, P.d.ya.p, P.qx.p, P.qy.p, P.rx.p,
                                                         P.ry.p);
    const int gc = std::max(1, (n + kThreads - 1) / kThreads), gr = std::max(1, (m + kThreads - 1) / kThreads);
    const CsrShape& AT = P.AT->shape;
    const CsrShape& A = P.A->shape;
    k_cert_cols<<<gc, kThreads, 0, s>>>(P.st.p, n, AT.nseg, AT.rowptr.p, AT.col.p, P.AT->val.p, P.ry.p, P.rx.p,
                                        P.d.c.p, P.d.lo.p, P.d.up.p, P.cert_c.p, split_len(AT));
    k_cert_rows<<<gr, kThreads, 0, s>>>(P.st.p, m, A.nseg, A.rowptr.p, A.col.p, P.A->val.p, P.rx.p, P.ry.p, P.d.rl.p,
                                        P.d.ru.p, P.cert_r.p, split_len(A));
    // Split rows: dd products per piece, then their terms into partial slots
    // after the one-lane kernels' slots.
    int nbc = gc, nbr = gr;
    if (AT.nsplit > 0) {
        k_cert_pieces<<<AT.nblk[0], kThreads, 0, s>>>(P.st.p, AT.view<double>(P.AT->val.p, 0), P.ry.p, P.cert_gt.p);
        const int gl = (AT.nsplit + kThreads - 1) / kThreads;
        k_cert_split_cols<<<gl, kThreads, 0, s>>>(P.st.p, AT.lrow.p, AT.nsplit, P.cert_gt.p, P.rx.p, P.d.c.p,
                                                  P.d.lo.p, P.d.up.p, P.cert_c.p + 6 * static_cast<size_t>(gc));
        nbc += gl;
    }
    if (A.nsplit > 0) {
        k_cert_pieces<<<A.nblk[0], kThreads, 0, s>>>(P.st.p, A.view<double>(P.A->val.p, 0), P.rx.p, P.cert_ga.p);
        const int gl = (A.nsplit + kThreads - 1) / kThreads;
        k_cert_split_rows<<<gl, kThreads, 0, s>>>(P.st.p, A.lrow.p, A.nsplit, P.cert_ga.p, P.ry.p, P.d.rl.p,
                                                  P.d.ru.p, P.cert_r.p + 6 * static_cast<size_t>(gr));
        nbr += gl;
    }
    k_cert_decide<<<1, kThreads, 0, s>>>(P.st.p, P.cert_c.p, nbc, P.cert_r.p, nbr);
}

template <typename TI>
void launch_eval(Pdlp& P, hipStream_t s, Iterates<TI>& it, int mode) {
    const int force = mode == 1 ? 1 : 0;
    const DevMatrix& AT = *P.AT;
    const DevMatrix& A = *P.A;
    // FP32 iterates are deltas from the FP64 anchor (X, Y) with products
    // GA = A^T Y, HA = A X.
    constexpr bool kDelta = std::is_same<TI, float>::value;
    const double* xanc = kDelta ? P.d.xa.p : nullptr;
    const double* yanc = kDelta ? P.d.ya.p : nullptr;
    const double* gbase = kDelta ? P.GA.p : nullptr;
    const double* hbase = kDelta ? P.HA.p : nullptr;
    for (int q = 0; q < AT.shape.nseg; ++q) {
        const Pass<double> ps = AT.shape.pass(AT.val.p, AT.part.p, q);
        TILE_DISPATCH(AT.shape.shape.tile,
                      k_eval_cols<TL, TI><<<ps.M.nblk, kThreads, 0, s>>>(ps, it.yp.p, it.xp.p, it.xa.p, P.d.c.p,
                                                                         P.d.lo.p, P.d.up.p, P.Ctot.p, xanc, gbase,
                                                                         P.st.p, P.part_c.p, force));
    }
    for (int q = 0; q < A.shape.nseg; ++q) {
        const Pass<double> ps = A.shape.pass(A.val.p, A.part.p, q);
        TILE_DISPATCH(A.shape.shape.tile,
                      k_eval_rows<TL, TI><<<ps.M.nblk, kThreads, 0, s>>>(ps, it.xp.p, it.yp.p, it.ya.p, P.d.rl.p,
                                                                         P.d.ru.p, P.Rtot.p, yanc, hbase, P.st.p,
                                                                         P.part_r.p, force));
    }
    const int nbc = AT.shape.nblk[AT.shape.nseg - 1];
    const int nbr = A.shape.nblk[A.shape.nseg - 1];
    k_decide<<<1, kThreads, 0, s>>>(P.st.p, P.part_c.p, nbc, P.part_r.p, nbr, mode);
    if (mode == 0) {
        // The FP64 anchor is d.xa / d.ya in both phases.
        const int gv = grid_for(std::max(P.n, P.m));
        k_cert_save<<<gv, 256, 0, s>>>(P.st.p, P.n, P.m, P.d.xa.p, P.d.ya.p, P.qx.p, P.qy.p);
        if (kDelta)
            recenter(P, s, 0);
        else
            k_restart_copy<TI><<<gv, 256, 0, s>>>(P.st.p, P.n, P.m, it.xp.p, it.yp.p, it.x.p, it.xa.p, it.y.p,
                                                  it.ya.p, 0);
        launch_certificate(P, s);
    }
    PD_CHECK(hipGetLastError());
}

template <typename T>
void enqueue_block(Pdlp& P, hipStream_t s, Iterates<T>& it) {
    for (int kl = 0; kl < P.eval_freq; ++kl) launch_iter<T>(P, s, it, kl, kl == P.eval_freq - 1 ? 1 : 0);
    launch_eval<T>(P, s, it, 0);
}

void enqueue(Pdlp& P, hipStream_t s) {
    if (P.use_graph) {
        if (!P.gexec || P.graph_mixed != P.mixed) {
            if (P.gexec) (void)hipGraphExecDestroy(P.gexec);
            if (P.graph) (void)hipGraphDestroy(P.graph);
            P.gexec = nullptr;
            P.graph = nullptr;
            PD_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            if (P.mixed) enqueue_block<float>(P, s, P.f);
            else enqueue_block<double>(P, s, P.d);
            PD_CHECK(hipStreamEndCapture(s, &P.graph));
            PD_CHECK(hipGraphInstantiate(&P.gexec, P.graph, nullptr, nullptr, 0));
            P.graph_mixed = P.mixed;
        }
        PD_CHECK(hipGraphLaunch(P.gexec, s));
    } else {
        if (P.mixed) enqueue_block<float>(P, s, P.f);
        else enqueue_block<double>(P, s, P.d);
    }
}

void post_poll(Pdlp& P, hipStream_t s, int slot) {
    PD_CHECK(hipMemcpyAsync(&P.h_st[slot], P.st.p, sizeof(PdlpState), hipMemcpyDeviceToHost, s));
    PD_CHECK(hipEventRecord(P.ev[slot], s));
    P.pending[slot] = 1;
}

bool wait_poll(Pdlp& P, int slot) {
    if (!P.pending[slot]) return false;
    PD_CHECK(hipEventSynchronize(P.ev[slot]));
    P.pending[slot] = 0;
    return P.h_st[slot].done != 0;
}

PdlpState read_state(Pdlp& P, hipStream_t s) {
    PdlpState h{};
    PD_CHECK(hipStreamSynchronize(s));
    P.pending[0] = P.pending[1] = 0;
    PD_CHECK(hipMemcpy(&h, P.st.p, sizeof(PdlpState), hipMemcpyDeviceToHost));
    return h;
}

void write_state(Pdlp& P, hipStream_t s, const PdlpState& h) {
    PD_CHECK(hipMemcpyAsync(P.st.p, &h, sizeof(PdlpState), hipMemcpyHostToDevice, s));
    PD_CHECK(hipStreamSynchronize(s));
}

// FP32 -> FP64 hand-over: the FP64 phase restarts at T(z) of the FP32 phase.
void switch_to_fp64(Pdlp& P, hipStream_t s, double new_tol) {
    const long long n = P.n, m = P.m;
    k_delta_to_double<<<grid_for(n), 256, 0, s>>>(P.d.xa.p, P.f.xp.p, P.d.xp.p, n);
    k_delta_to_double<<<grid_for(m), 256, 0, s>>>(P.d.ya.p, P.f.yp.p, P.d.yp.p, m);
    k_restart_copy<double><<<grid_for(std::max(n, m)), 256, 0, s>>>(P.st.p, P.n, P.m, P.d.xp.p, P.d.yp.p, P.d.x.p,
                                                                    P.d.xa.p, P.d.y.p, P.d.ya.p, 1);
    PdlpState h = read_state(P, s);
    h.done = 0;
    h.status = 0;
    h.k_base = 0;
    h.tol = new_tol;
    h.watch_stall = 0;
    write_state(P, s, h);
    P.mixed = false;
    P.fp64_iters_at_switch = h.total;
    // Anchor KKT of the hand-over point, in FP64.
    launch_eval<double>(P, s, P.d, 1);
    PD_CHECK(hipStreamSynchronize(s));
}

void spmv_plain(const DevMatrix& M, const double* in, double* out, hipStream_t s) {
    for (int q = 0; q < M.shape.nseg; ++q) {
        const Pass<double> ps = M.shape.pass(M.val.p, M.part.p, q);
        TILE_DISPATCH(M.shape.shape.tile, k_spmv_plain<TL><<<ps.M.nblk, kThreads, 0, s>>>(ps, in, out));
    }
    PD_CHECK(hipGetLastError());
}

}  // namespace

double pdlp_prepare_matrix(hipStream_t s, int m, int n, long long nnz, DevMatrix* Am, DevMatrix* ATm, double* dR,
                           double* dC, std::vector<double>& R, std::vector<double>& Cc) {
    // ---- exponent-domain Ruiz (10 passes) + Pock-Chambolle, on the device
    Buf<double> mr, mc, fr, fc, onesr, onesc;
    mr.alloc(m); mc.alloc(n); fr.alloc(m); fc.alloc(n); onesr.alloc(m); onesc.alloc(n);
    std::vector<double> ones_m(static_cast<size_t>(m), 1.0), ones_n(static_cast<size_t>(n), 1.0);
    if (m) PD_CHECK(hipMemcpyAsync(dR, ones_m.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
    if (n) PD_CHECK(hipMemcpyAsync(dC, ones_n.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    onesr.up(ones_m.data(), m, s);
    onesc.up(ones_n.data(), n, s);
    const CsrShape& SA = Am->shape;
    const CsrShape& ST = ATm->shape;
    double* a_val = Am->val.p;
    double* at_val = ATm->val.p;
    const int gm = (m + 255) / 256, gn = (n + 255) / 256;
    // One lane per row, and one workgroup per piece of a split row.
    auto measure = [&](const CsrShape& S, double* val, int rows, double* out, int what) {
        k_row_measure<<<(rows + 255) / 256, 256, 0, s>>>(S.rowptr.p, S.nseg, rows, val, out, what, split_len(S));
        if (S.nsplit > 0) {
            k_piece_measure<<<S.nblk[0], kThreads, 0, s>>>(S.view<double>(val, 0), out, what);
            if (what == 1) k_split_total<<<(S.nsplit + 255) / 256, 256, 0, s>>>(S.lrow.p, S.nsplit, S.gsum.p, out);
        }
    };
    auto scale = [&](const CsrShape& S, double* val, int rows, const double* rf, const double* cf) {
        k_scale_csr<<<(rows + 255) / 256, 256, 0, s>>>(S.rowptr.p, S.nseg, rows, S.col.p, val, rf, cf, split_len(S));
        if (S.nsplit > 0) k_piece_scale<<<S.nblk[0], kThreads, 0, s>>>(S.view<double>(val, 0), val, rf, cf);
    };
    for (int pass = 0; pass < 11; ++pass) {
        const int what = pass < 10 ? 0 : 1;   // 10 Ruiz (max |a|) passes, then one Pock-Chambolle (sum |a|)
        if (m > 0) {
            measure(SA, a_val, m, mr.p, what);
            k_factor<<<gm, 256, 0, s>>>(mr.p, fr.p, dR, m);
            scale(SA, a_val, m, fr.p, onesc.p);
        }
        if (n > 0) {
            scale(ST, at_val, n, onesc.p, fr.p);
            measure(ST, at_val, n, mc.p, what);
            k_factor<<<gn, 256, 0, s>>>(mc.p, fc.p, dC, n);
            scale(ST, at_val, n, fc.p, onesr.p);
        }
        if (m > 0) scale(SA, a_val, m, onesr.p, fc.p);
        PD_CHECK(hipGetLastError());
    }
    R.assign(static_cast<size_t>(m), 0.0);
    Cc.assign(static_cast<size_t>(n), 0.0);
    if (m) PD_CHECK(hipMemcpyAsync(R.data(), dR, m * sizeof(double), hipMemcpyDeviceToHost, s));
    if (n) PD_CHECK(hipMemcpyAsync(Cc.data(), dC, n * sizeof(double), hipMemcpyDeviceToHost, s));
    PD_CHECK(hipStreamSynchronize(s));

    // ---- power iteration on A'^T A'
    Buf<double> u, v;
    u.alloc(n);
    v.alloc(m);
    {
        std::vector<double> u0(n);
        for (int j = 0; j < n; ++j) u0[j] = 1.0 + static_cast<double>((static_cast<unsigned>(j) * 2654435761u) % 1000u) / 1000.0;
        u.up(u0.data(), n, s);
    }
    Buf<dd> sq;
    const int nsq = std::max(1, std::min(1024, (n + 255) / 256));
    sq.alloc(nsq);
    std::vector<dd> hsq(nsq);
    auto norm_of_u = [&]() {
        k_sumsq<<<nsq, kThreads, 0, s>>>(u.p, n, sq.p);
        sq.down(hsq.data(), nsq, s);
        PD_CHECK(hipStreamSynchronize(s));
        dd t = dd_zero();
        for (int b = 0; b < nsq; ++b) t = dd_add(t, hsq[b]);
        return std::sqrt(dd_val(t));
    };
    double lam = 0.0;
    if (n > 0 && m > 0 && nnz > 0) {
        double nu = norm_of_u();
        if (nu > 0) k_scale_vec<<<gn, 256, 0, s>>>(u.p, n, nu);
        for (int it = 0; it < 200; ++it) {
            spmv_plain(*Am, u.p, v.p, s);
            spmv_plain(*ATm, v.p, u.p, s);
            const double nw = norm_of_u();
            if (!(nw > 0.0) || !std::isfinite(nw)) break;
            k_scale_vec<<<gn, 256, 0, s>>>(u.p, n, nw);
            const double prev = lam;
            lam = nw;
            if (it > 0 && std::fabs(lam - prev) <= 1e-7 * lam) break;
        }
    }
    PD_CHECK(hipStreamSynchronize(s));
    const double smax = lam > 0.0 ? std::sqrt(lam) : 1.0;
    return 0.998 / smax;
}

void pdlp_prepare_instance(const nir_lp* lp, const nir_params& prm, const std::vector<double>& R,
                           const std::vector<double>& Cc, double eta, bool mixed, PdlpInstance& out) {
    const int n = lp->ncol, m = lp->nrow;
    const double tol = prm.tol > 0 ? prm.tol : 1e-6;
    // ---- scaled vectors (absent bounds as +-inf) and caller-unit scales
    auto finite_b = [](double b) { return std::isfinite(b) && std::fabs(b) < 1e20; };
    std::vector<double>&c = out.c, &lo = out.lo, &up = out.up, &rl = out.rl, &ru = out.ru, &rt = out.rt,
                       &ct = out.ct;
    c.resize(n); lo.resize(n); up.resize(n); rl.resize(m); ru.resize(m); rt.resize(m); ct.resize(n);
    DDHost bn, cn;
    for (int j = 0; j < n; ++j) {
        c[j] = lp->cost[j] * Cc[j];
        lo[j] = finite_b(lp->collo[j]) ? lp->collo[j] / Cc[j] : -INFINITY;
        up[j] = finite_b(lp->colup[j]) ? lp->colup[j] / Cc[j] : INFINITY;
        const double us = prm.unscale_col ? prm.unscale_col[j] : 1.0;
        ct[j] = Cc[j] / us;   // lambda_caller = lambda' / C  ->  lambda_user = lambda_caller / unscale
        const double cu = lp->cost[j] / us;
        cn.add_prod(cu, cu);
    }
    for (int i = 0; i < m; ++i) {
        rl[i] = finite_b(lp->rowlo[i]) ? lp->rowlo[i] * R[i] : -INFINITY;
        ru[i] = finite_b(lp->rowup[i]) ? lp->rowup[i] * R[i] : INFINITY;
        const double us = prm.unscale_row ? prm.unscale_row[i] : 1.0;
        rt[i] = R[i] * us;
        double b = 0.0;
        if (finite_b(lp->rowlo[i])) b = std::fabs(lp->rowlo[i]);
        if (finite_b(lp->rowup[i])) b = std::max(b, std::fabs(lp->rowup[i]));
        b = b / us;
        bn.add_prod(b, b);
    }
    // ---- start point: clamp(0) or the warm start, in the scaled space
    out.x0.assign(static_cast<size_t>(n), 0.0);
    out.y0.assign(static_cast<size_t>(m), 0.0);
    for (int j = 0; j < n; ++j) {
        double v = prm.x0 ? prm.x0[j] / Cc[j] : 0.0;
        v = (lo[j] < v) ? v : lo[j];
        v = (v < up[j]) ? v : up[j];
        out.x0[j] = v;
    }
    if (prm.y0)
        for (int i = 0; i < m; ++i) out.y0[i] = prm.y0[i] / R[i];
    // ---- primal weight and state
    DDHost cs, bs;
    for (int j = 0; j < n; ++j) cs.add_prod(c[j], c[j]);
    for (int i = 0; i < m; ++i) {
        double b = 0.0;
        if (std::isfinite(rl[i])) b = std::fabs(rl[i]);
        if (std::isfinite(ru[i])) b = std::max(b, std::fabs(ru[i]));
        bs.add_prod(b, b);
    }
    const double cnorm = std::sqrt(cs.value()), bnorm = std::sqrt(bs.value());
    const double omega = prm.primal_weight > 0 ? prm.primal_weight
                                               : ((cnorm > 1e-10 && bnorm > 1e-10) ? cnorm / bnorm : 1.0);
    PdlpState& h = out.st;
    h = PdlpState{};
    h.omega = omega;
    h.eta = eta;
    h.tau = eta / omega;
    h.sigma = eta * omega;
    h.tol = mixed ? tol : (prm.polish_tol > 0 ? std::min(tol, prm.polish_tol) : tol);
    h.bnorm_o = std::sqrt(bn.value());
    h.cnorm_o = std::sqrt(cn.value());
    h.offset = lp->obj_offset;
    h.max_iters = std::max<long long>(1, prm.max_iters);
    h.eval_freq = std::max(1, prm.eval_freq);
    h.watch_stall = mixed ? 1 : 0;
}

// Poll slots, events, evaluation partials and certificate buffers.
static void alloc_work(Pdlp& P) {
    const int n = P.n, m = P.m;
    const CsrShape& SA = P.A->shape;
    const CsrShape& ST = P.AT->shape;
    PD_CHECK(hipHostMalloc(reinterpret_cast<void**>(&P.h_st), 2 * sizeof(PdlpState), hipHostMallocDefault));
    std::memset(P.h_st, 0, 2 * sizeof(PdlpState));
    for (auto& e : P.ev) PD_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    P.part_c.alloc(static_cast<size_t>(5) * ST.nblk[ST.nseg - 1]);
    P.part_r.alloc(static_cast<size_t>(6) * SA.nblk[SA.nseg - 1]);
    P.qx.alloc(n);
    P.qy.alloc(m);
    P.rx.alloc(n);
    P.ry.alloc(m);
    const auto blocks = [](int k) { return (k + kThreads - 1) / kThreads; };
    P.cert_c.alloc(static_cast<size_t>(6) * (std::max(1, blocks(n)) + blocks(ST.nsplit)));
    P.cert_r.alloc(static_cast<size_t>(6) * (std::max(1, blocks(m)) + blocks(SA.nsplit)));
    P.cert_gt.alloc(std::max(1, ST.ngroups));
    P.cert_ga.alloc(std::max(1, SA.ngroups));
}

void pdlp_setup(Pdlp& P, hipStream_t s, const nir_lp* lp, const nir_params& prm, DevMatrix* Am, DevMatrix* ATm) {
    auto t0 = std::chrono::steady_clock::now();
    P.n = lp->ncol;
    P.m = lp->nrow;
    P.nnz = P.n > 0 ? lp->col_start[P.n] : 0;
    P.A = Am;
    P.AT = ATm;
    P.eval_freq = std::max(1, prm.eval_freq);
    P.use_graph = prm.use_graph != 0;
    P.tol = prm.tol > 0 ? prm.tol : 1e-6;
    P.mixed = prm.precision == NIR_PREC_MIXED;
    P.polish_tol = prm.polish_tol;
    P.offset = lp->obj_offset;
    P.should_stop = prm.should_stop;
    P.stop_user = prm.stop_user;
    P.h_c.assign(lp->cost, lp->cost + P.n);
    const int n = P.n, m = P.m;
    double* a_val = Am->val.p;
    double* at_val = ATm->val.p;

    P.R.alloc(m);
    P.C.alloc(n);
    std::vector<double> R, Cc;
    const double eta = pdlp_prepare_matrix(s, m, n, P.nnz, Am, ATm, P.R.p, P.C.p, R, Cc);
    PdlpInstance I;
    pdlp_prepare_instance(lp, prm, R, Cc, eta, P.mixed, I);
    P.d.alloc(n, m);
    P.d.c.up(I.c.data(), n, s);
    P.d.lo.up(I.lo.data(), n, s);
    P.d.up.up(I.up.data(), n, s);
    P.d.rl.up(I.rl.data(), m, s);
    P.d.ru.up(I.ru.data(), m, s);
    P.Rtot.alloc(m);
    P.Ctot.alloc(n);
    P.Rtot.up(I.rt.data(), m, s);
    P.Ctot.up(I.ct.data(), n, s);
    for (auto* b : {&P.d.x, &P.d.xa, &P.d.xp, &P.d.xbar}) b->up(I.x0.data(), n, s);
    for (auto* b : {&P.d.y, &P.d.ya, &P.d.yp}) b->up(I.y0.data(), m, s);
    const PdlpState h = I.st;
    P.st.alloc(1);
    P.st.up(&h, 1, s);
    alloc_work(P);

    if (P.mixed) {
        P.a32.val.alloc(P.nnz);
        P.at32.val.alloc(P.nnz);
        P.a32.part.alloc(std::max(m, 1));
        P.at32.part.alloc(std::max(n, 1));
        k_to_float<<<grid_for(P.nnz), 256, 0, s>>>(a_val, P.a32.val.p, P.nnz);
        k_to_float<<<grid_for(P.nnz), 256, 0, s>>>(at_val, P.at32.val.p, P.nnz);
        P.f.alloc(n, m);
        P.HA.alloc(m);
        P.GA.alloc(n);
        // The anchor is the start point (d.xa / d.ya); deltas start at zero
        // (k_recenter writes them together with the re-based constants).
        recenter(P, s, 1);
        launch_eval<float>(P, s, P.f, 1);
    } else {
        launch_eval<double>(P, s, P.d, 1);
    }
    PdlpState hs = read_state(P, s);
    if (hs.done) {
        P.finished = hs.status != 0 || !(P.mixed && P.polish_tol > 0);
        P.status = hs.status;
        if (!P.finished && P.mixed) {
            P.polish_done = true;
            switch_to_fp64(P, s, P.polish_tol);
            PdlpState h2 = read_state(P, s);
            if (h2.done) {
                P.finished = true;
                P.status = h2.status;
            }
        }
    }
    P.setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// A finished state: the end of the run, or, for an FP32 phase that met its
// tolerance with a polish to do or stalled, the switch to the FP64 tail.
// Returns true when the run goes on (in FP64).
static bool on_done(Pdlp& P, hipStream_t s, const PdlpState& h) {
    if (P.mixed && (h.status == 0 || h.status == 2) && !P.polish_done) {
        const double target = P.polish_tol > 0 ? std::min(P.polish_tol, P.tol) : P.tol;
        if (h.status == 0 && P.polish_tol <= 0) {
            P.finished = true;
            P.status = 0;
            return false;
        }
        P.polish_done = true;
        switch_to_fp64(P, s, target);
        PdlpState h2 = read_state(P, s);
        if (h2.done) {
            P.finished = true;
            P.status = sov_status(h2);
            return false;
        }
        return true;
    }
    P.finished = true;
    P.status = sov_status(h);
    return false;
}

void pdlp_adopt(Pdlp& P, hipStream_t s, const nir_params& prm, const PdlpAdopt& a) {
    auto t0 = std::chrono::steady_clock::now();
    P.n = a.n;
    P.m = a.m;
    P.nnz = a.nnz;
    P.A = a.A;
    P.AT = a.AT;
    P.eval_freq = std::max(1, prm.eval_freq);
    P.use_graph = prm.use_graph != 0;
    P.tol = prm.tol > 0 ? prm.tol : 1e-6;
    P.mixed = a.mixed;
    P.polish_tol = prm.polish_tol;
    P.offset = a.offset;
    P.should_stop = prm.should_stop;
    P.stop_user = prm.stop_user;
    P.h_c = a.h_c;
    const int n = P.n, m = P.m;
    P.R.alloc(m);
    P.C.alloc(n);
    P.R.from_device(a.R, m, s);
    P.C.from_device(a.C, n, s);
    // No unscaling vectors on this path: the KKT units are R and C.
    P.Rtot.alloc(m);
    P.Ctot.alloc(n);
    P.Rtot.from_device(a.R, m, s);
    P.Ctot.from_device(a.C, n, s);
    P.d.alloc(n, m);
    P.d.c.from_device(a.c, n, s);
    P.d.lo.from_device(a.lo, n, s);
    P.d.up.from_device(a.up, n, s);
    P.d.rl.from_device(a.rl, m, s);
    P.d.ru.from_device(a.ru, m, s);
    P.d.x.from_device(a.x, n, s);
    P.d.xbar.from_device(a.xbar, n, s);
    P.d.xa.from_device(a.xa, n, s);
    P.d.xp.from_device(a.xp, n, s);
    P.d.y.from_device(a.y, m, s);
    P.d.ya.from_device(a.ya, m, s);
    P.d.yp.from_device(a.yp, m, s);
    P.st.alloc(1);
    P.st.from_device(a.st, 1, s);
    alloc_work(P);
    if (P.mixed) {
        P.a32.val.alloc(P.nnz);
        P.at32.val.alloc(P.nnz);
        P.a32.part.alloc(std::max(m, 1));
        P.at32.part.alloc(std::max(n, 1));
        k_to_float<<<grid_for(P.nnz), 256, 0, s>>>(P.A->val.p, P.a32.val.p, P.nnz);
        k_to_float<<<grid_for(P.nnz), 256, 0, s>>>(P.AT->val.p, P.at32.val.p, P.nnz);
        PD_CHECK(hipGetLastError());
        P.f.alloc(n, m);
        P.HA.alloc(m);
        P.GA.alloc(n);
        P.HA.from_device(a.HA, m, s);
        P.GA.from_device(a.GA, n, s);
        P.f.x.from_device(a.fx, n, s);
        P.f.xbar.from_device(a.fxbar, n, s);
        P.f.xp.from_device(a.fxp, n, s);
        P.f.c.from_device(a.rc, n, s);
        P.f.lo.from_device(a.dl, n, s);
        P.f.up.from_device(a.du, n, s);
        P.f.y.from_device(a.fy, m, s);
        P.f.yp.from_device(a.fyp, m, s);
        P.f.rl.from_device(a.drl, m, s);
        P.f.ru.from_device(a.dru, m, s);
        P.f.ya.from_device(a.sa, m, s);
    }
    const PdlpState h = read_state(P, s);
    if (h.done) on_done(P, s, h);
    P.setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void pdlp_run(Pdlp& P, hipStream_t s, const nir_params& prm, long long budget) {
    auto t0 = std::chrono::steady_clock::now();
    long long blocks = (budget + P.eval_freq - 1) / P.eval_freq;
    const long long poll_every = std::max(1, 256 / P.eval_freq);
    long long since_poll = 0;
    while (!P.finished && blocks > 0) {
        if ((P.polls % poll_every) == 0 && P.should_stop && P.should_stop(P.stop_user)) {
            PdlpState h = read_state(P, s);
            P.status = 4;
            P.finished = true;
            h.done = 1;
            h.status = 4;
            write_state(P, s, h);
            break;
        }
        ++P.polls;
        enqueue(P, s);
        --blocks;
        if (++since_poll >= poll_every || blocks == 0) {
            since_poll = 0;
            const int slot = static_cast<int>(P.serial++ & 1);
            post_poll(P, s, slot);
            if (wait_poll(P, slot ^ 1) || blocks == 0) {
                PdlpState h = read_state(P, s);
                if (h.done) {
                    // FP32 phase met its tolerance (or stalled): FP64 tail.
                    if (on_done(P, s, h)) continue;
                    break;
                }
            }
        }
    }
    PD_CHECK(hipStreamSynchronize(s));
    P.solve_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int pdlp_result(Pdlp& P, hipStream_t s, double* x, double* y, nir_info* info) {
    PdlpState h = read_state(P, s);
    std::vector<double> xs(static_cast<size_t>(P.n)), ys(static_cast<size_t>(P.m));
    if (P.mixed) {
        // T(z) of the FP32 phase: anchor + delta in FP64.
        std::vector<float> xf(static_cast<size_t>(P.n)), yf(static_cast<size_t>(P.m));
        P.f.xp.down(xf.data(), P.n, s);
        P.f.yp.down(yf.data(), P.m, s);
        P.d.xa.down(xs.data(), P.n, s);
        P.d.ya.down(ys.data(), P.m, s);
        PD_CHECK(hipStreamSynchronize(s));
        for (int j = 0; j < P.n; ++j) xs[j] = xs[j] + static_cast<double>(xf[j]);
        for (int i = 0; i < P.m; ++i) ys[i] = ys[i] + static_cast<double>(yf[i]);
    } else {
        P.d.xp.down(xs.data(), P.n, s);
        P.d.yp.down(ys.data(), P.m, s);
    }
    std::vector<double> R(static_cast<size_t>(P.m)), Cc(static_cast<size_t>(P.n));
    P.R.down(R.data(), P.m, s);
    P.C.down(Cc.data(), P.n, s);
    PD_CHECK(hipStreamSynchronize(s));
    for (int j = 0; j < P.n; ++j) xs[j] *= Cc[j];
    for (int i = 0; i < P.m; ++i) ys[i] *= R[i];
    if (x) std::copy(xs.begin(), xs.end(), x);
    if (y) std::copy(ys.begin(), ys.end(), y);
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->status = P.finished ? P.status : (h.done ? sov_status(h) : 6);
        info->iterations = h.total;
        info->polish_iterations = P.fp64_iters_at_switch >= 0 ? h.total - P.fp64_iters_at_switch : 0;
        info->restarts = h.restarts;
        DDHost acc(P.offset);
        for (int j = 0; j < P.n; ++j)
            if (P.h_c[j] != 0.0 && xs[j] != 0.0) acc.add_prod(P.h_c[j], xs[j]);
        info->objective = acc.value();
        info->tau = h.tau;
        info->sigma = h.sigma;
        info->primal_weight = h.omega;
        info->rel_primal = h.rel_p;
        info->rel_dual = h.rel_d;
        info->rel_gap = h.rel_g;
        info->setup_seconds = P.setup_s;
        info->solve_seconds = P.solve_s;
    }
    return 0;
}

// The accumulated Ruiz + Pock-Chambolle factors (R: m, C: n; either may be NULL).
void pdlp_scaling(Pdlp& P, hipStream_t s, double* R, double* C) {
    if (R) P.R.down(R, static_cast<size_t>(P.m), s);
    if (C) P.C.down(C, static_cast<size_t>(P.n), s);
    PD_CHECK(hipStreamSynchronize(s));
}

int pdlp_time(Pdlp& P, hipStream_t s, long long iters, nir_timing* out) {
    // Events around every iteration kernel of eval_freq-blocks (all segment
    // passes of a product between one pair of events).
    const long long blocks = std::max<long long>(1, iters / P.eval_freq);
    const int n = P.n, m = P.m;
    const double nnz = static_cast<double>(P.nnz);
    const double sv = P.mixed ? 4.0 : 8.0;
    // A xbar + dual step: values+cols, rowptr, xbar once, y/ya/rl/ru read, y write.
    out->spmv_bytes = nnz * (sv + 4) + 4.0 * (m + 1) + sv * n + sv * 5 * m;
    // A^T y + primal step: values+cols, rowptr, y once, x/xa/c/lo/up read, x/xbar write
    // (the re-centred FP32 form reads no anchor).
    out->spmtv_bytes = nnz * (sv + 4) + 4.0 * (n + 1) + sv * m + sv * (P.mixed ? 6 : 7) * n;
    out->iter_bytes = out->spmv_bytes + out->spmtv_bytes;
    out->spmv_passes = P.A->shape.nseg;
    out->spmtv_passes = P.AT->shape.nseg;
    std::vector<hipEvent_t> ev(static_cast<size_t>(4 * blocks * P.eval_freq + 2));
    for (auto& e : ev) PD_CHECK(hipEventCreate(&e));
    PD_CHECK(hipStreamSynchronize(s));
    PD_CHECK(hipEventRecord(ev[0], s));
    size_t q = 2;
    for (long long b = 0; b < blocks; ++b) {
        for (int kl = 0; kl < P.eval_freq; ++kl) {
            const int save = kl == P.eval_freq - 1 ? 1 : 0;
            PD_CHECK(hipEventRecord(ev[q++], s));
            if (P.mixed) launch_primal<float>(P, s, P.f, kl, save);
            else launch_primal<double>(P, s, P.d, kl, save);
            PD_CHECK(hipEventRecord(ev[q++], s));
            PD_CHECK(hipEventRecord(ev[q++], s));
            if (P.mixed) launch_dual<float>(P, s, P.f, kl, save);
            else launch_dual<double>(P, s, P.d, kl, save);
            PD_CHECK(hipEventRecord(ev[q++], s));
        }
        if (P.mixed) launch_eval<float>(P, s, P.f, 0);
        else launch_eval<double>(P, s, P.d, 0);
    }
    PD_CHECK(hipEventRecord(ev[1], s));
    PD_CHECK(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    PD_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    out->total_ms = ms;
    for (size_t k = 2; k + 3 < q; k += 4) {
        PD_CHECK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
        out->spmtv_ms += ms;
        PD_CHECK(hipEventElapsedTime(&ms, ev[k + 2], ev[k + 3]));
        out->spmv_ms += ms;
    }
    for (auto& e : ev) PD_CHECK(hipEventDestroy(e));
    out->iterations = blocks * P.eval_freq;
    return 0;
}

}  // namespace gpu
}  // namespace nir
