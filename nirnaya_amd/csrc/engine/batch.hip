// Batched restarted PDHG: many LPs of uneven size in persistent launches.
//
// Layout in HBM: every instance's CSR(A) and CSR(A^T) (int32 local indices,
// FP64 values scaled in place), their wave-sliced copies (FP64, and FP32 in
// mixed mode) and its vectors are concatenated; an instance table holds the
// offsets. An instance's rows and columns are stored by decreasing length
// (stable orders; every row of A and of A^T keeps its entries in the
// caller's order under the new labels), so that the 64-row slices of the
// row loops hold rows of like length; inputs and results go through the
// orders on the host, and the power iteration's start vector is a function
// of the caller's column index (colmap). One workgroup owns one instance at
// a time and runs whole PDHG iterations on it with workgroup barriers
// between the primal and the dual half.
//
// Operand staging: the products gather y (for A^T y) and xbar (for A xbar)
// at random. When the largest instance's max(m, n) FP64 entries fit in one
// workgroup's LDS (up to 152 KB on gfx950), the 1024-lane kernel keeps the
// operand being gathered in LDS, so the random reads hit LDS instead of the
// vector memory path: an instance whose m + n entries fit (in the precision
// of the phase) keeps y and xbar in two regions that the lanes computing
// them write directly, y staying there across an evaluation block; a larger
// one keeps one region - y on entry, xbar after the primal half, y again
// after the dual half, each copied in from global memory after a barrier.
//
// Matrix streams: the row loops read wave-sliced copies of A and A^T (see
// slice_dot), so a wave's value and index loads are contiguous runs; their
// column indices take 16 bits when every instance has max(m, n) <= 65,536
// (the planning family does), halving the index stream.
//
// Work queue: a list of unfinished instances, largest first. A launch of the
// persistent kernel hands them out by an atomic ticket; the workgroup runs up
// to `quantum` evaluation blocks (eval_freq iterations + FP64 evaluation +
// restart decision each) on its instance, writes the instance's state back
// and takes the next ticket; it exits when the list is exhausted, so every
// launch drains. A one-workgroup kernel then compacts the list (finished
// instances leave it) and re-arms the ticket counter. The host replays a
// hipGraph of several (run, compact) rounds and polls the list length.
//
// Every per-instance computation uses the single-LP engine's arithmetic
// (pdlp_core.hpp): sequential row and column sums in the caller's index
// order, the same scaling, power iteration, step sizes, Halpern updates,
// evaluation terms and decisions; norms are double-double (their rounding
// to double does not depend on the order of the terms, the device row
// order included, short of a sum within about 2^-100 of a rounding
// boundary). So each instance's iterates, iteration and restart counts
// equal what nir_pdhg_solve returns for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../../include/nirnaya/pdhg.h"
#include "ctx.hpp"
#include "devmat.hpp"
#include "lpcheck.hpp"
#include "pdlp.hpp"
#include "pdlp_core.hpp"
#include "trace.hpp"

namespace nir {
namespace gpu {
namespace batch {

// Offsets of one instance in the concatenated arrays.
struct Inst {
    int m, n, nnz, pad;
    long long arp;    // CSR(A) row pointers (m + 1)
    long long trp;    // CSR(A^T) row pointers (n + 1)
    long long nz;     // nonzeros (both matrices)
    long long xo;     // column vectors
    long long yo;     // row vectors
};

// Evaluation blocks per work-list ticket by default: large enough that most
// instances run to completion on their first ticket, so the largest-first
// list schedules whole solves longest first (time-slicing instances with
// short tickets measured 30 % slower on 512 planning LPs: 2.81 s at 16 blocks,
// 1.96 s at 4096).
constexpr int kDefaultQuantum = 1024;

// Per-instance control beyond the iteration state.
struct Ctl {
    int fp32;          // 1 while the instance iterates in FP32 (mixed mode)
    int finished;
    int status;
    int polish_done;
    long long fp64_at; // iterations completed at the FP64 hand-over (-1: none)
};

template <typename T>
struct Vec {
    T *x, *xbar, *xa, *xp, *y, *ya, *yp;
    T *c, *lo, *up, *rl, *ru;
    T *aval, *tval;   // CSR values (FP64 only; scaled in place at setup)
    T *sa, *st;       // the same values wave-sliced (see slice_dot)
};

struct View {
    const Inst* inst;
    const int *arow, *acol, *trow, *tcol;
    void *sacol, *stcol;  // wave-sliced column indices of A and A^T (written at setup),
    int idx16;            // uint16 when every instance has max(m, n) <= 65536, else int32
    Vec<double> d;
    Vec<float> f;
    double *R, *C, *hs;
    double *HA, *GA;   // FP32 phase: A X and A^T Y of the FP64 anchor (X, Y) = (d.xa, d.ya)
    const double *cost, *collo, *colup, *rowlo, *rowup, *offset;
    PdlpState* st;
    Ctl* ctl;
    int* list;
    int* count;
    int* head;
    long long* launches;
    int mixed;
    int eval_freq;
    int quantum;
    double tol, polish_tol;
    double omega0;     // initial primal weight (0: from the norms)
    long long max_iters;
    int stage_bytes;   // dynamic LDS of the launch for staged operands (0: none)
    const double* x0;  // warm starts (concatenated like the vectors), with
    const double* y0;  // per-instance flags; null when no instance has one
    const int* has_x0;
    const int* has_y0;
    const int* colmap;  // device column j of an instance is the caller's column colmap[xo + j]
};

// Instance-local pointers of one precision.
template <typename T>
struct Local {
    const int *arow, *acol, *trow, *tcol;
    const void *sacol, *stcol;   // uint16 (idx16) or int32
    int idx16;
    T *aval, *tval, *sa, *st;
    T *x, *xbar, *xa, *xp, *y, *ya, *yp;
    T *c, *lo, *up, *rl, *ru;
};

template <typename T>
__device__ __forceinline__ Local<T> local(const View& V, const Vec<T>& v, const Inst& I) {
    Local<T> L;
    L.arow = V.arow + I.arp;
    L.acol = V.acol + I.nz;
    L.trow = V.trow + I.trp;
    L.tcol = V.tcol + I.nz;
    const long long ib = I.nz * (V.idx16 ? 2 : 4);
    L.sacol = static_cast<const char*>(V.sacol) + ib;
    L.stcol = static_cast<const char*>(V.stcol) + ib;
    L.idx16 = V.idx16;
    L.aval = v.aval ? v.aval + I.nz : nullptr;   // CSR values: FP64 only
    L.tval = v.tval ? v.tval + I.nz : nullptr;
    L.sa = v.sa + I.nz;
    L.st = v.st + I.nz;
    L.x = v.x + I.xo; L.xbar = v.xbar + I.xo; L.xa = v.xa + I.xo; L.xp = v.xp + I.xo;
    L.c = v.c + I.xo; L.lo = v.lo + I.xo; L.up = v.up + I.xo;
    L.y = v.y + I.yo; L.ya = v.ya + I.yo; L.yp = v.yp + I.yo;
    L.rl = v.rl + I.yo; L.ru = v.ru + I.yo;
    return L;
}

// Wave-sliced copies of the CSR matrices for the row loops. An instance's
// rows are cut into slices of 64: the rows one wave takes in every row loop
// of this file (row = threadIdx.x + k * BT, BT a multiple of 64). A slice
// keeps the CSR range of its rows but orders it by position within the row
// first: the first nonzeros of all its nonempty rows in row order, then the
// second nonzeros of the rows longer than one, and so on. At step t a wave
// loads one contiguous run of values and indices instead of 64 strided ones
// (with one lane per CSR row, each lane's loads are a stream of their own),
// and every row still adds its products in CSR order, one rounded product at
// a time: the sequential row sums of the single-LP CSR-stream kernels, bit
// for bit. A call must come from a full row loop: the wave's active lanes
// are exactly the slice's rows below the row count, lane 0 holding the
// slice's first row.
__device__ __forceinline__ int lane_rank(unsigned long long mask) {
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u)));
}

// base[i] addressed by a 32-bit byte offset from a uniform base (the
// scalar-base form of the load, without 64-bit address arithmetic per lane;
// per-instance arrays are below 2^28 entries, checked at batch creation).
template <typename T>
__device__ __forceinline__ T at32(const T* __restrict__ base, unsigned i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * static_cast<unsigned>(sizeof(T)));
}

template <typename T, typename IT, typename V>
__device__ __forceinline__ T slice_dot(const int* __restrict__ rp, const IT* __restrict__ scol,
                                       const T* __restrict__ sval, const V* __restrict__ v, int r) {
    const int start = rp[r];
    const int len = rp[r + 1] - start;
    int base = __builtin_amdgcn_readfirstlane(start);   // the slice's CSR start
    T s = T(0);
    // U steps per round: one workgroup owns an instance, so the loop is
    // bound by memory latency, not bandwidth; all U steps' loads and then
    // all U gathers are issued before the first product is added. Eight in
    // FP32 (most rows of the planning and staircase families take one
    // round), four in FP64, where eight value and operand pairs cost more
    // registers than the extra requests in flight gain. Lanes past their
    // row's end load from `base` (in bounds: step t has an active lane) and
    // keep their sum by a select.
    constexpr int U = sizeof(T) == 4 ? 8 : 4;
    for (int t = 0;; t += U) {
        unsigned long long mk[U];
        mk[0] = __ballot(t < len);
        if (mk[0] == 0ull) break;
#pragma unroll
        for (int u = 1; u < U; ++u) mk[u] = __ballot(t + u < len);
        int p[U];
        int b = base;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            p[u] = t + u < len ? b + lane_rank(mk[u]) : base;
            b += __popcll(mk[u]);
        }
        // The active steps of a round are a prefix (a row active at step
        // t + 1 is active at t): steps past it are skipped by the whole wave.
        int nact = 1;
#pragma unroll
        for (int u = 1; u < U; ++u) nact += mk[u] != 0ull ? 1 : 0;
        T a[U];
        int c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (u < nact) {
                a[u] = at32(sval, static_cast<unsigned>(p[u]));
                c[u] = static_cast<int>(at32(scol, static_cast<unsigned>(p[u])));
            }
        }
        T g[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (u < nact) g[u] = static_cast<T>(at32(v, static_cast<unsigned>(c[u])));
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (u < nact) {
                const T su = s + a[u] * g[u];
                s = t + u < len ? su : s;
            }
        }
        base = b;
    }
    return s;
}

// slice_dot over the batch's index width (uniform: one width per batch).
template <typename T, typename V>
__device__ __forceinline__ T slice_row(const int* __restrict__ rp, const void* scol, int idx16,
                                       const T* __restrict__ sval, const V* __restrict__ v, int r) {
    return idx16 ? slice_dot<T>(rp, static_cast<const unsigned short*>(scol), sval, v, r)
                 : slice_dot<T>(rp, static_cast<const int*>(scol), sval, v, r);
}

// Writes the wave-sliced copy of a CSR matrix (rows x ?, row pointers rp):
// the positions slice_dot reads, by the same traversal.
template <int BT, typename T>
__device__ void wg_slice(const int* __restrict__ rp, const int* __restrict__ col, const T* __restrict__ val,
                         int rows, void* scol, int idx16, T* __restrict__ sval) {
    for (int r = threadIdx.x; r < rows; r += BT) {
        const int start = rp[r];
        const int len = rp[r + 1] - start;
        int base = __builtin_amdgcn_readfirstlane(start);
        for (int t = 0;; ++t) {
            const unsigned long long mask = __ballot(t < len);
            if (mask == 0ull) break;
            if (t < len) {
                const int p = base + lane_rank(mask);
                if (idx16) static_cast<unsigned short*>(scol)[p] = static_cast<unsigned short>(col[start + t]);
                else static_cast<int*>(scol)[p] = col[start + t];
                sval[p] = val[start + t];
            }
            base += __popcll(mask);
        }
    }
}

// dd sums of NQ per-thread accumulators over the workgroup (wave butterflies,
// then the waves in order); every thread receives the totals in `tot`.
template <int NQ, int BT>
__device__ __forceinline__ void wg_dd_sum(dd (&acc)[NQ], dd* sh, dd* tot) {
    constexpr int NW = BT / 64;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] = dd_add(acc[q], dd{__shfl_xor(acc[q].hi, o), __shfl_xor(acc[q].lo, o)});
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) sh[q * NW + w] = acc[q];
    __syncthreads();
    if (threadIdx.x < NQ) {
        dd t = dd_zero();
        for (int k = 0; k < NW; ++k) t = dd_add(t, sh[threadIdx.x * NW + k]);
        tot[threadIdx.x] = t;
    }
    __syncthreads();
}

// Workgroup-shared scratch of the persistent kernels.
template <int BT>
struct Shared {
    PdlpState s;
    Ctl ctl;
    dd red[11 * (BT / 64)];
    dd tot[11];
    double scalar;
    int item;
};

// ------------------------------------------------------------------ evaluation

// FP64 evaluation of T(z) (xp, yp in precision TI) and the decision.
template <typename TI, int BT>
__device__ void wg_eval(const View& V, const Inst& I, const Vec<TI>& it, Shared<BT>& sh, int mode) {
    const Local<double> D = local(V, V.d, I);
    const TI* xp = it.xp + I.xo;
    const TI* xa = it.xa + I.xo;
    const TI* yp = it.yp + I.yo;
    const TI* ya = it.ya + I.yo;
    const double* C = V.C + I.xo;
    const double* R = V.R + I.yo;
    // Products first (xbar is dead between blocks and holds A^T yp), then the
    // terms: two light loops instead of one that keeps gathers and eleven
    // double-double accumulators live together.
    double* gs = V.d.xbar + I.xo;
    double* hs = V.hs + I.yo;
    for (int j = threadIdx.x; j < I.n; j += BT) gs[j] = slice_row<double>(D.trow, D.stcol, D.idx16, D.st, yp, j);
    for (int i = threadIdx.x; i < I.m; i += BT) hs[i] = slice_row<double>(D.arow, D.sacol, D.idx16, D.sa, xp, i);
    // FP32 iterates are deltas from the FP64 anchor: T(z) = anchor + delta,
    // A^T yp = GA + A^T dyp, A xp = HA + A dxp.
    constexpr bool kDelta = std::is_same<TI, float>::value;
    const double* GA = V.GA + I.xo;
    const double* HA = V.HA + I.yo;
    {
        dd acc[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) acc[q] = dd_zero();
        for (int j = threadIdx.x; j < I.n; j += BT) {
            dd t[5];
            if (kDelta)
                col_terms(GA[j] + gs[j], D.c[j], D.lo[j], D.up[j], C[j], D.xa[j] + static_cast<double>(xp[j]),
                          D.xa[j], t);
            else
                col_terms(gs[j], D.c[j], D.lo[j], D.up[j], C[j], static_cast<double>(xp[j]),
                          static_cast<double>(xa[j]), t);
#pragma unroll
            for (int q = 0; q < 5; ++q) acc[q] = dd_add(acc[q], t[q]);
        }
        wg_dd_sum<5, BT>(acc, sh.red, sh.tot);
    }
    {
        dd acc[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[q] = dd_zero();
        for (int i = threadIdx.x; i < I.m; i += BT) {
            dd t[6];
            if (kDelta)
                row_terms(HA[i] + hs[i], D.rl[i], D.ru[i], R[i], D.ya[i] + static_cast<double>(yp[i]), D.ya[i], t);
            else
                row_terms(hs[i], D.rl[i], D.ru[i], R[i], static_cast<double>(yp[i]), static_cast<double>(ya[i]), t);
#pragma unroll
            for (int q = 0; q < 6; ++q) acc[q] = dd_add(acc[q], t[q]);
        }
        wg_dd_sum<6, BT>(acc, sh.red, sh.tot + 5);
    }
    if (threadIdx.x == 0) pdlp_decide(sh.s, sh.tot, mode);
    __syncthreads();
}

// ------------------------------------------------------------------ iteration

template <typename T, int BT>
__device__ void wg_iteration(const Local<T>& L, const Inst& I, const PdlpState& s, int k_lo, int save) {
    const Halpern<T> hw = halpern<T>(s.k_base + k_lo);
    const T tau = static_cast<T>(s.tau);
    constexpr bool kDelta = std::is_same<T, float>::value;   // FP32: re-centred form, zero anchor delta
    for (int j = threadIdx.x; j < I.n; j += BT) {
        const T xj = L.x[j], cj = L.c[j], lj = L.lo[j], uj = L.up[j], aj = kDelta ? T(0) : L.xa[j];
        const T g = slice_row<T>(L.trow, L.stcol, L.idx16, L.st, L.y, j);
        const PrimalOut<T> r = primal_point<T>(xj, g, cj, lj, uj, aj, tau, hw);
        L.xbar[j] = r.xb;
        L.x[j] = r.next;
        if (save) L.xp[j] = r.cand;
    }
    __syncthreads();
    const T sigma = static_cast<T>(s.sigma);
    for (int i = threadIdx.x; i < I.m; i += BT) {
        const T yi = L.y[i], rli = L.rl[i], rui = L.ru[i], ai = L.ya[i];
        const T h = slice_row<T>(L.arow, L.sacol, L.idx16, L.sa, L.xbar, i);
        const DualOut<T> r = kDelta ? dual_point_delta<T>(yi, h, rli, rui, ai, sigma, hw)
                                    : dual_point<T>(yi, h, rli, rui, ai, sigma, hw);
        L.y[i] = r.next;
        if (save) L.yp[i] = r.yp;
    }
    __syncthreads();
}

// The launch's dynamic LDS, where an instance's gathered operand is staged.
template <typename T>
__device__ __forceinline__ T* stage_base() {
    extern __shared__ double batch_stage[];
    return reinterpret_cast<T*>(batch_stage);
}

// LDS copy of `count` entries written earlier by other lanes, four loads in
// flight per lane before the LDS writes (the copy is latency-bound).
template <typename T, int BT>
__device__ __forceinline__ void stage_copy(T* dst, const T* __restrict__ src, int count) {
    for (int k = threadIdx.x; k < count; k += 4 * BT) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = k + u * BT < count ? src[k + u * BT] : T(0);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + u * BT < count) dst[k + u * BT] = v[u];
    }
}

// The same iteration with the gathered operands in LDS. The products read
// the same values in the same order, so the iterates are bit-identical to
// the unstaged form; the random accesses go to LDS instead of the vector
// memory path. `stage` holds y on entry and on return.
//   two regions (m + n entries fit): y at stage[0, m), xbar at stage[m, m+n);
//     each lane writes the xbar and y entries it computes straight into
//     LDS (nothing reads the region a half writes until the barrier after
//     it), so an iteration has two barriers and no copies; y is read from
//     its region and written to global memory only by the block's last
//     iteration (`save`), where the evaluation and the next block read it;
//   one region (max(m, n) entries): y, then xbar after the primal half, then
//     y again, each copied in from global memory after a barrier.
template <typename T, int BT>
__device__ __forceinline__ void wg_iteration_staged(const Local<T>& L, const Inst& I, const PdlpState& s, int k_lo,
                                                    int save, T* stage, bool two) {
    const Halpern<T> hw = halpern<T>(s.k_base + k_lo);
    const T tau = static_cast<T>(s.tau);
    constexpr bool kDelta = std::is_same<T, float>::value;
    T* sx = two ? stage + I.m : stage;
    for (int j = threadIdx.x; j < I.n; j += BT) {
        const T xj = L.x[j], cj = L.c[j], lj = L.lo[j], uj = L.up[j], aj = kDelta ? T(0) : L.xa[j];
        const T g = slice_row<T>(L.trow, L.stcol, L.idx16, L.st, stage, j);
        const PrimalOut<T> r = primal_point<T>(xj, g, cj, lj, uj, aj, tau, hw);
        if (two) sx[j] = r.xb;
        else L.xbar[j] = r.xb;
        L.x[j] = r.next;
        if (save) L.xp[j] = r.cand;
    }
    __syncthreads();
    if (!two) {
        stage_copy<T, BT>(stage, L.xbar, I.n);
        __syncthreads();
    }
    const T sigma = static_cast<T>(s.sigma);
    for (int i = threadIdx.x; i < I.m; i += BT) {
        const T yi = two ? stage[i] : L.y[i], rli = L.rl[i], rui = L.ru[i], ai = L.ya[i];
        const T h = slice_row<T>(L.arow, L.sacol, L.idx16, L.sa, sx, i);
        const DualOut<T> r = kDelta ? dual_point_delta<T>(yi, h, rli, rui, ai, sigma, hw)
                                    : dual_point<T>(yi, h, rli, rui, ai, sigma, hw);
        if (two) stage[i] = r.next;
        if (!two || save) L.y[i] = r.next;
        if (save) L.yp[i] = r.yp;
    }
    __syncthreads();
    if (!two) {
        stage_copy<T, BT>(stage, L.y, I.m);
        __syncthreads();
    }
}

// Restart: iterate and anchor become T(z).
template <typename T, int BT>
__device__ __forceinline__ void wg_restart_copy(const Local<T>& L, const Inst& I) {
    for (int j = threadIdx.x; j < I.n; j += BT) {
        const T v = L.xp[j];
        L.x[j] = v;
        L.xa[j] = v;
    }
    for (int i = threadIdx.x; i < I.m; i += BT) {
        const T v = L.yp[i];
        L.y[i] = v;
        L.ya[i] = v;
    }
    __syncthreads();
}

// Re-centre the FP32 phase: the anchor moves to T(z) (skipped at setup,
// `force`), A X and A^T Y are recomputed in FP64 and the FP32 constants are
// re-based on the anchor with zero deltas (the single-LP recenter sequence).
template <int BT>
__device__ void wg_recenter(const View& V, const Inst& I, const PdlpState& s, int force) {
    const Local<double> D = local(V, V.d, I);
    const Local<float> F = local(V, V.f, I);
    double* HA = V.HA + I.yo;
    double* GA = V.GA + I.xo;
    if (!force) {
        for (int j = threadIdx.x; j < I.n; j += BT) D.xa[j] = D.xa[j] + static_cast<double>(F.xp[j]);
        for (int i = threadIdx.x; i < I.m; i += BT) D.ya[i] = D.ya[i] + static_cast<double>(F.yp[i]);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < I.m; i += BT) HA[i] = slice_row<double>(D.arow, D.sacol, D.idx16, D.sa, D.xa, i);
    for (int j = threadIdx.x; j < I.n; j += BT) GA[j] = slice_row<double>(D.trow, D.stcol, D.idx16, D.st, D.ya, j);
    __syncthreads();
    const double sigma = s.sigma;
    for (int j = threadIdx.x; j < I.n; j += BT) {
        F.c[j] = static_cast<float>(D.c[j] - GA[j]);
        F.lo[j] = static_cast<float>(D.lo[j] - D.xa[j]);
        F.up[j] = static_cast<float>(D.up[j] - D.xa[j]);
        F.x[j] = 0.0f;
        F.xp[j] = 0.0f;
    }
    for (int i = threadIdx.x; i < I.m; i += BT) {
        F.rl[i] = static_cast<float>(D.rl[i] - HA[i]);
        F.ru[i] = static_cast<float>(D.ru[i] - HA[i]);
        F.ya[i] = static_cast<float>(-D.ya[i] / sigma);
        F.y[i] = 0.0f;
        F.yp[i] = 0.0f;
    }
    __syncthreads();
}

// FP32 -> FP64 hand-over: the FP64 phase restarts at T(z) of the FP32 phase.
template <int BT>
__device__ void wg_switch_to_fp64(const View& V, const Inst& I, Shared<BT>& sh, double new_tol) {
    const Local<double> D = local(V, V.d, I);
    const float* fxp = V.f.xp + I.xo;
    const float* fyp = V.f.yp + I.yo;
    for (int j = threadIdx.x; j < I.n; j += BT) D.xp[j] = D.xa[j] + static_cast<double>(fxp[j]);
    for (int i = threadIdx.x; i < I.m; i += BT) D.yp[i] = D.ya[i] + static_cast<double>(fyp[i]);
    __syncthreads();
    wg_restart_copy<double, BT>(D, I);
    if (threadIdx.x == 0) {
        sh.s.done = 0;
        sh.s.status = 0;
        sh.s.k_base = 0;
        sh.s.tol = new_tol;
        sh.s.watch_stall = 0;
        sh.ctl.fp32 = 0;
        sh.ctl.fp64_at = sh.s.total;
    }
    __syncthreads();
    wg_eval<double, BT>(V, I, V.d, sh, 1);
}

// What follows a decision that set done (the host loop of the single-LP
// engine, made per instance): finish, or hand the FP32 phase over to FP64.
// st, fp32 and polish_done are every thread's own copies of the shared
// state, taken before a barrier: thread 0 rewrites that state below, so no
// thread may branch on the shared copy after this point.
template <int BT>
__device__ void wg_after_done(const View& V, const Inst& I, Shared<BT>& sh, int st, int fp32, int polish_done) {
    if (fp32 && (st == 0 || st == 2) && !polish_done) {
        const double target = V.polish_tol > 0 ? fmin(V.polish_tol, V.tol) : V.tol;
        if (st == 0 && V.polish_tol <= 0) {
            if (threadIdx.x == 0) {
                sh.ctl.finished = 1;
                sh.ctl.status = 0;
            }
            __syncthreads();
            return;
        }
        if (threadIdx.x == 0) sh.ctl.polish_done = 1;
        __syncthreads();
        wg_switch_to_fp64<BT>(V, I, sh, target);
        if (threadIdx.x == 0 && sh.s.done) {
            sh.ctl.finished = 1;
            sh.ctl.status = sh.s.status;
        }
        __syncthreads();
        return;
    }
    if (threadIdx.x == 0) {
        sh.ctl.finished = 1;
        sh.ctl.status = st;
    }
    __syncthreads();
}

// Infeasibility certificate of a restart (terms and decision in
// pdlp_core.hpp), taken before the anchor moves. The step between the anchors
// is formed as the single-LP engine forms it after the move: T(z) - anchor in
// FP64, (X + dxp) - X in the FP32 phase. Returns the same code for every
// thread (0, kStPrimalInfeasible or kStDualInfeasible).
template <typename T, int BT>
__device__ int wg_certificate(const View& V, const Inst& I, Shared<BT>& sh) {
    const Local<double> D = local(V, V.d, I);
    const float* fxp = V.f.xp + I.xo;
    const float* fyp = V.f.yp + I.yo;
    constexpr bool kF = std::is_same<T, float>::value;
    auto sx = [&](int j) -> double {
        return kF ? (D.xa[j] + static_cast<double>(fxp[j])) - D.xa[j] : D.xp[j] - D.xa[j];
    };
    auto sy = [&](int i) -> double {
        return kF ? (D.ya[i] + static_cast<double>(fyp[i])) - D.ya[i] : D.yp[i] - D.ya[i];
    };
    dd acc[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] = dd_zero();
    for (int j = threadIdx.x; j < I.n; j += BT) {
        dd g = dd_zero();
        for (int p = D.trow[j]; p < D.trow[j + 1]; ++p) g = dd_add(g, dd_prod(D.tval[p], sy(D.tcol[p])));
        cert_col_terms(-dd_val(g), D.c[j], D.lo[j], D.up[j], sx(j), acc);
    }
    for (int i = threadIdx.x; i < I.m; i += BT) {
        dd h = dd_zero();
        for (int p = D.arow[i]; p < D.arow[i + 1]; ++p) h = dd_add(h, dd_prod(D.aval[p], sx(D.acol[p])));
        cert_row_terms(dd_val(h), D.rl[i], D.ru[i], sy(i), acc);
    }
    wg_dd_sum<6, BT>(acc, sh.red, sh.tot);
    const int code = cert_decide(sh.tot);
    __syncthreads();
    return code;
}

// One evaluation block: eval_freq iterations, evaluation, restart copy.
// STAGED: the launch stages every instance's gathered operands in LDS (the
// host launches it only when the largest instance fits).
template <typename T, int BT, bool STAGED>
__device__ void wg_block(const View& V, const Inst& I, const Vec<T>& v, Shared<BT>& sh) {
    const Local<T> L = local(V, v, I);
    if (STAGED) {
        T* stage = stage_base<T>();
        const bool two = static_cast<long long>(I.m + I.n) * static_cast<long long>(sizeof(T)) <= V.stage_bytes;
        stage_copy<T, BT>(stage, L.y, I.m);
        __syncthreads();
        for (int kl = 0; kl < V.eval_freq; ++kl)
            wg_iteration_staged<T, BT>(L, I, sh.s, kl, kl == V.eval_freq - 1 ? 1 : 0, stage, two);
    } else {
        for (int kl = 0; kl < V.eval_freq; ++kl) wg_iteration<T, BT>(L, I, sh.s, kl, kl == V.eval_freq - 1 ? 1 : 0);
    }
    wg_eval<T, BT>(V, I, v, sh, 0);
    // Decisions into registers, then a barrier: every thread branches on the
    // same values although thread 0 rewrites the shared state afterwards.
    const int restart = sh.s.restart, done = sh.s.done, st = sh.s.status;
    const int fp32 = sh.ctl.fp32, polish_done = sh.ctl.polish_done;
    __syncthreads();
    if (restart) {
        const int cert = wg_certificate<T, BT>(V, I, sh);
        if (std::is_same<T, float>::value) wg_recenter<BT>(V, I, sh.s, 0);
        else wg_restart_copy<T, BT>(L, I);
        if (cert) {
            if (threadIdx.x == 0) {
                sh.s.done = 1;
                sh.s.status = cert;
                sh.s.restart = 0;
            }
            __syncthreads();
            wg_after_done<BT>(V, I, sh, cert, fp32, polish_done);
        }
    }
    if (done) wg_after_done<BT>(V, I, sh, st, fp32, polish_done);
}

// Take the next ticket of the work list; -1 when the list is exhausted.
template <int BT>
__device__ __forceinline__ int wg_pop(const View& V, Shared<BT>& sh) {
    if (threadIdx.x == 0) {
        const int t = atomicAdd(V.head, 1);
        sh.item = t < *V.count ? V.list[t] : -1;
    }
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(sh.item);   // uniform: instance pointers stay scalar
    __syncthreads();
    return i;
}

// 4 waves per SIMD: the register budget of the iteration loop (the rare
// evaluation code spills a few dwords).
template <int BT, bool STAGED>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_batch_run(View V) {
    __shared__ Shared<BT> sh;
    for (;;) {
        const int i = wg_pop<BT>(V, sh);
        if (i < 0) return;
        const Inst I = V.inst[i];
        if (threadIdx.x == 0) {
            sh.s = V.st[i];
            sh.ctl = V.ctl[i];
        }
        __syncthreads();
        for (int b = 0; b < V.quantum; ++b) {
            const int finished = sh.ctl.finished, fp32 = sh.ctl.fp32;
            __syncthreads();
            if (finished) break;
            if (fp32) wg_block<float, BT, STAGED>(V, I, V.f, sh);
            else wg_block<double, BT, STAGED>(V, I, V.d, sh);
        }
        if (threadIdx.x == 0) {
            V.st[i] = sh.s;
            V.ctl[i] = sh.ctl;
        }
        __syncthreads();
    }
}

// Drop finished instances from the list (order kept) and re-arm the tickets.
__global__ void k_batch_compact(View V) {
    if (threadIdx.x != 0) return;
    const int cnt = *V.count;
    int k = 0;
    for (int t = 0; t < cnt; ++t) {
        const int i = V.list[t];
        if (!V.ctl[i].finished) V.list[k++] = i;
    }
    *V.count = k;
    *V.head = 0;
    *V.launches += 1;
}

// ------------------------------------------------------------------ setup

// Sum of squares (double-double) of v over the workgroup.
template <int BT>
__device__ double wg_norm(const double* v, int n, Shared<BT>& sh) {
    dd acc[1] = {dd_zero()};
    for (int j = threadIdx.x; j < n; j += BT) acc[0] = dd_add(acc[0], dd_prod(v[j], v[j]));
    wg_dd_sum<1, BT>(acc, sh.red, sh.tot);
    return sqrt(dd_val(sh.tot[0]));
}

// Row measure: max |a| (what 0) or sum |a| (what 1), ascending index order.
__device__ __forceinline__ double row_measure(const int* rp, const double* val, int r, int what) {
    double acc = 0.0;
    for (int p = rp[r]; p < rp[r + 1]; ++p) {
        const double a = fabs(val[p]);
        acc = what == 0 ? amax(acc, a) : acc + a;
    }
    return acc;
}

// Scaling (10 Ruiz passes, one Pock-Chambolle pass; powers of two), scaled
// vectors, start point, power iteration, primal weight, FP32 copies and the
// evaluation of the start point, for one instance.
template <int BT>
__device__ void wg_setup(const View& V, int i, Shared<BT>& sh) {
    const Inst I = V.inst[i];
    const int m = I.m, n = I.n, tid = threadIdx.x;
    const Local<double> D = local(V, V.d, I);
    double* aval = D.aval;
    double* tval = D.tval;
    double* R = V.R + I.yo;
    double* C = V.C + I.xo;
    double* fr = D.ya;   // scratch until the start point is written
    double* fc = D.xa;
    for (int r = tid; r < m; r += BT) R[r] = 1.0;
    for (int j = tid; j < n; j += BT) C[j] = 1.0;
    __syncthreads();
    for (int pass = 0; pass < 11; ++pass) {
        const int what = pass < 10 ? 0 : 1;
        if (m > 0) {
            for (int r = tid; r < m; r += BT) {
                const double g = pow2_inv_sqrt(row_measure(D.arow, aval, r, what));
                fr[r] = g;
                R[r] *= g;
            }
            __syncthreads();
            for (int r = tid; r < m; r += BT)
                for (int p = D.arow[r]; p < D.arow[r + 1]; ++p) aval[p] = aval[p] * (fr[r] * 1.0);
        }
        if (n > 0) {
            for (int j = tid; j < n; j += BT)
                for (int p = D.trow[j]; p < D.trow[j + 1]; ++p) tval[p] = tval[p] * (1.0 * fr[D.tcol[p]]);
            __syncthreads();
            for (int j = tid; j < n; j += BT) {
                const double g = pow2_inv_sqrt(row_measure(D.trow, tval, j, what));
                fc[j] = g;
                C[j] *= g;
            }
            __syncthreads();
            for (int j = tid; j < n; j += BT)
                for (int p = D.trow[j]; p < D.trow[j + 1]; ++p) tval[p] = tval[p] * (fc[j] * 1.0);
        }
        if (m > 0)
            for (int r = tid; r < m; r += BT)
                for (int p = D.arow[r]; p < D.arow[r + 1]; ++p) aval[p] = aval[p] * (1.0 * fc[D.acol[p]]);
        __syncthreads();
    }
    // Wave-sliced copies of the scaled matrices for the row loops.
    wg_slice<BT>(D.arow, D.acol, aval, m, const_cast<void*>(D.sacol), V.idx16, D.sa);
    wg_slice<BT>(D.trow, D.tcol, tval, n, const_cast<void*>(D.stcol), V.idx16, D.st);
    __syncthreads();

    // Scaled vectors (absent bounds as +-inf) and the norms.
    const double* cost = V.cost + I.xo;
    const double* collo = V.collo + I.xo;
    const double* colup = V.colup + I.xo;
    const double* rowlo = V.rowlo + I.yo;
    const double* rowup = V.rowup + I.yo;
    dd acc[4] = {dd_zero(), dd_zero(), dd_zero(), dd_zero()};   // caller |c|^2, |b|^2; scaled |c|^2, |b|^2
    for (int j = tid; j < n; j += BT) {
        const double cj = cost[j] * C[j];
        D.c[j] = cj;
        D.lo[j] = finite_bound(collo[j]) ? collo[j] / C[j] : -INFINITY;
        D.up[j] = finite_bound(colup[j]) ? colup[j] / C[j] : INFINITY;
        acc[0] = dd_add(acc[0], dd_prod(cost[j], cost[j]));
        acc[2] = dd_add(acc[2], dd_prod(cj, cj));
    }
    for (int r = tid; r < m; r += BT) {
        const double rl = finite_bound(rowlo[r]) ? rowlo[r] * R[r] : -INFINITY;
        const double ru = finite_bound(rowup[r]) ? rowup[r] * R[r] : INFINITY;
        D.rl[r] = rl;
        D.ru[r] = ru;
        double b = 0.0;
        if (finite_bound(rowlo[r])) b = fabs(rowlo[r]);
        if (finite_bound(rowup[r])) b = amax(b, fabs(rowup[r]));
        acc[1] = dd_add(acc[1], dd_prod(b, b));
        double bs = 0.0;
        if (isfinite(rl)) bs = fabs(rl);
        if (isfinite(ru)) bs = amax(bs, fabs(ru));
        acc[3] = dd_add(acc[3], dd_prod(bs, bs));
    }
    wg_dd_sum<4, BT>(acc, sh.red, sh.tot);
    const double cnorm_o = sqrt(dd_val(sh.tot[0])), bnorm_o = sqrt(dd_val(sh.tot[1]));
    const double cnorm = sqrt(dd_val(sh.tot[2])), bnorm = sqrt(dd_val(sh.tot[3]));
    __syncthreads();

    // Power iteration on A'^T A' (u in xbar, v in yp as scratch).
    double lam = 0.0;
    if (n > 0 && m > 0 && I.nnz > 0) {
        double* u = D.xbar;
        double* v = D.yp;
        for (int j = tid; j < n; j += BT) {
            const unsigned jo = static_cast<unsigned>(V.colmap[I.xo + j]);   // the caller's column index
            u[j] = 1.0 + static_cast<double>((jo * 2654435761u) % 1000u) / 1000.0;
        }
        __syncthreads();
        const double nu = wg_norm<BT>(u, n, sh);
        if (nu > 0) {
            for (int j = tid; j < n; j += BT) u[j] = u[j] / nu;
            __syncthreads();
        }
        for (int it = 0; it < 200; ++it) {
            for (int r = tid; r < m; r += BT) v[r] = slice_row<double>(D.arow, D.sacol, D.idx16, D.sa, u, r);
            __syncthreads();
            for (int j = tid; j < n; j += BT) u[j] = slice_row<double>(D.trow, D.stcol, D.idx16, D.st, v, j);
            __syncthreads();
            const double nw = wg_norm<BT>(u, n, sh);
            if (!(nw > 0.0) || !isfinite(nw)) break;
            for (int j = tid; j < n; j += BT) u[j] = u[j] / nw;
            __syncthreads();
            const double prev = lam;
            lam = nw;
            if (it > 0 && fabs(lam - prev) <= 1e-7 * lam) break;
        }
    }
    const double smax = lam > 0.0 ? sqrt(lam) : 1.0;
    const double eta = 0.998 / smax;

    // Start point: the warm start (in caller units) or 0, clamped into the
    // bounds; y = the warm start or 0.
    const bool wx = V.x0 && V.has_x0[i], wy = V.y0 && V.has_y0[i];
    for (int j = tid; j < n; j += BT) {
        double v = wx ? V.x0[I.xo + j] / C[j] : 0.0;
        v = (D.lo[j] < v) ? v : D.lo[j];
        v = (v < D.up[j]) ? v : D.up[j];
        D.x[j] = v;
        D.xa[j] = v;
        D.xp[j] = v;
        D.xbar[j] = v;
    }
    for (int r = tid; r < m; r += BT) {
        const double v = wy ? V.y0[I.yo + r] / R[r] : 0.0;
        D.y[r] = v;
        D.ya[r] = v;
        D.yp[r] = v;
    }
    if (tid == 0) {
        PdlpState h{};
        h.omega = V.omega0 > 0 ? V.omega0 : ((cnorm > 1e-10 && bnorm > 1e-10) ? cnorm / bnorm : 1.0);
        h.eta = eta;
        h.tau = eta / h.omega;
        h.sigma = eta * h.omega;
        h.tol = V.mixed ? V.tol : (V.polish_tol > 0 ? fmin(V.tol, V.polish_tol) : V.tol);
        h.bnorm_o = bnorm_o;
        h.cnorm_o = cnorm_o;
        h.offset = V.offset[i];
        h.max_iters = V.max_iters;
        h.eval_freq = V.eval_freq;
        h.watch_stall = V.mixed ? 1 : 0;
        sh.s = h;
        Ctl c{};
        c.fp32 = V.mixed;
        c.fp64_at = -1;
        sh.ctl = c;
    }
    __syncthreads();

    if (V.mixed) {
        const Local<float> F = local(V, V.f, I);
        for (int p = tid; p < I.nnz; p += BT) {
            F.sa[p] = static_cast<float>(D.sa[p]);
            F.st[p] = static_cast<float>(D.st[p]);
        }
        __syncthreads();
        // The anchor is the start point (D.xa / D.ya); re-base the FP32
        // constants on it with zero deltas.
        wg_recenter<BT>(V, I, sh.s, 1);
        __syncthreads();
        wg_eval<float, BT>(V, I, V.f, sh, 1);
    } else {
        wg_eval<double, BT>(V, I, V.d, sh, 1);
    }
    // The single-LP setup's handling of a start point that already meets the
    // tolerance (finish, or polish from it in FP64).
    const int done0 = sh.s.done, st = sh.s.status;
    __syncthreads();
    if (done0) {
        const bool fin = st != 0 || !(V.mixed && V.polish_tol > 0);
        if (threadIdx.x == 0) {
            sh.ctl.finished = fin ? 1 : 0;
            sh.ctl.status = st;
        }
        __syncthreads();
        if (!fin && V.mixed) {
            if (threadIdx.x == 0) sh.ctl.polish_done = 1;
            __syncthreads();
            wg_switch_to_fp64<BT>(V, I, sh, V.polish_tol);
            if (threadIdx.x == 0 && sh.s.done) {
                sh.ctl.finished = 1;
                sh.ctl.status = sh.s.status;
            }
            __syncthreads();
        }
    }
    if (tid == 0) {
        V.st[i] = sh.s;
        V.ctl[i] = sh.ctl;
    }
    __syncthreads();
}

template <int BT>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_batch_setup(View V) {
    __shared__ Shared<BT> sh;
    for (;;) {
        const int i = wg_pop<BT>(V, sh);
        if (i < 0) return;
        wg_setup<BT>(V, i, sh);
    }
}

// ------------------------------------------------------------------ host

template <typename T>
struct HostVec {
    DevArray<T> x, xbar, xa, xp, y, ya, yp, c, lo, up, rl, ru, aval, tval, sa, st;
    // csr: also the CSR value arrays (the FP64 set: scaling and certificates).
    void alloc(size_t nx, size_t ny, size_t nz, bool csr) {
        for (auto* a : {&x, &xbar, &xa, &xp, &c, &lo, &up}) a->alloc(nx);
        for (auto* a : {&y, &ya, &yp, &rl, &ru}) a->alloc(ny);
        if (csr) {
            aval.alloc(nz);
            tval.alloc(nz);
        }
        sa.alloc(nz);
        st.alloc(nz);
    }
    Vec<T> view() {
        return Vec<T>{x.p, xbar.p, xa.p, xp.p, y.p, ya.p, yp.p, c.p, lo.p, up.p, rl.p, ru.p, aval.p, tval.p, sa.p, st.p};
    }
    // Arrays the iteration changes (autotune snapshots): iterates, and in the
    // re-centred FP32 phase also the re-based constants.
    std::vector<DevArray<T>*> mutable_arrays() { return {&x, &xbar, &xa, &xp, &y, &ya, &yp, &c, &lo, &up, &rl, &ru}; }
};

}  // namespace batch
}  // namespace gpu
}  // namespace nir

using namespace nir::gpu;
using namespace nir::gpu::batch;

struct nir_batch {
    nir_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    nir_params prm{};
    nir_batch_opts opts{};
    int count = 0;
    std::vector<Inst> h_inst;
    std::vector<std::vector<double>> h_cost;   // for the objective of the result
    std::vector<std::vector<int>> h_order;     // per instance: device row r is the caller's row h_order[r]
    std::vector<std::vector<int>> h_corder;    // per instance: device column j is the caller's column h_corder[j]
    std::vector<double> h_offset;
    DevArray<Inst> inst;
    DevArray<int> arow, acol, trow, tcol;
    DevArray<unsigned char> sacol, stcol;   // wave-sliced indices, 2 or 4 bytes each
    int idx16 = 0;
    HostVec<double> d;
    HostVec<float> f;
    DevArray<double> R, C, hs, HA, GA, cost, collo, colup, rowlo, rowup, offset, x0, y0;
    DevArray<int> has_x0, has_y0, colmap;
    bool warm = false;
    DevArray<PdlpState> st;
    DevArray<Ctl> ctl;
    DevArray<int> list, count_d, head;
    DevArray<long long> launches;
    int* h_poll = nullptr;   // pinned: list length after the last graph replay (two slots)
    hipEvent_t poll_ev[2] = {nullptr, nullptr};
    int threads = 512, workgroups = 0;
    int stage_bytes = 0;   // dynamic LDS of the 1024-lane run kernel (largest staged operand)
    bool solved = false, stopped = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double setup_s = 0.0, solve_s = 0.0, total_nnz = 0.0;
    long long total_x = 0, total_y = 0, total_nz = 0;

    View view() {
        View V{};
        V.inst = inst.p;
        V.arow = arow.p;
        V.acol = acol.p;
        V.trow = trow.p;
        V.tcol = tcol.p;
        V.sacol = sacol.p;
        V.stcol = stcol.p;
        V.idx16 = idx16;
        V.d = d.view();
        V.f = prm.precision == NIR_PREC_MIXED ? f.view() : Vec<float>{};
        V.R = R.p;
        V.C = C.p;
        V.hs = hs.p;
        V.HA = HA.p;
        V.GA = GA.p;
        V.cost = cost.p;
        V.collo = collo.p;
        V.colup = colup.p;
        V.rowlo = rowlo.p;
        V.rowup = rowup.p;
        V.offset = offset.p;
        V.st = st.p;
        V.ctl = ctl.p;
        V.list = list.p;
        V.count = count_d.p;
        V.head = head.p;
        V.launches = launches.p;
        V.mixed = prm.precision == NIR_PREC_MIXED ? 1 : 0;
        V.eval_freq = std::max(1, prm.eval_freq);
        V.quantum = opts.quantum > 0 ? opts.quantum : kDefaultQuantum;
        V.tol = prm.tol > 0 ? prm.tol : 1e-6;
        V.polish_tol = prm.polish_tol;
        V.omega0 = prm.primal_weight > 0 ? prm.primal_weight : 0.0;
        V.max_iters = std::max<long long>(1, prm.max_iters);
        V.x0 = warm ? x0.p : nullptr;
        V.y0 = warm ? y0.p : nullptr;
        V.has_x0 = warm ? has_x0.p : nullptr;
        V.has_y0 = warm ? has_y0.p : nullptr;
        V.colmap = colmap.p;
        return V;
    }
    ~nir_batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : poll_ev)
            if (e) (void)hipEventDestroy(e);
        if (h_poll) (void)hipHostFree(h_poll);
    }
};

namespace {

void launch_run(nir_batch* B, const View& V0, int threads) {
    // Operand staging in LDS with the 1024-lane workgroups (one per CU),
    // when the largest instance fits.
    View V = V0;
    V.stage_bytes = threads == 1024 ? B->stage_bytes : 0;
    switch (threads) {
        case 256: k_batch_run<256, false><<<B->workgroups, 256, 0, B->s>>>(V); break;
        case 1024:
            if (V.stage_bytes > 0) k_batch_run<1024, true><<<B->workgroups, 1024, V.stage_bytes, B->s>>>(V);
            else k_batch_run<1024, false><<<B->workgroups, 1024, 0, B->s>>>(V);
            break;
        default: k_batch_run<512, false><<<B->workgroups, 512, 0, B->s>>>(V); break;
    }
    k_batch_compact<<<1, 64, 0, B->s>>>(V);
    HIP_CHECK(hipGetLastError());
}

void launch_setup(nir_batch* B, const View& V, int threads) {
    switch (threads) {
        case 256: k_batch_setup<256><<<B->workgroups, 256, 0, B->s>>>(V); break;
        case 1024: k_batch_setup<1024><<<B->workgroups, 1024, 0, B->s>>>(V); break;
        default: k_batch_setup<512><<<B->workgroups, 512, 0, B->s>>>(V); break;
    }
    HIP_CHECK(hipGetLastError());
}

int compute_units(int device) {
    int cu = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device));
    return std::max(1, cu);
}

// Autotuner: time one quantum-1 launch of every workgroup size from the
// post-setup state, restoring the state after each; keep the fastest. The
// choice never changes results (each instance's arithmetic is the same).
int autotune_threads(nir_batch* B, int cu) {
    std::vector<int> cands = {256, 512, 1024};
    View V = B->view();
    V.quantum = 1;
    std::vector<DevArray<double>*> dm = B->d.mutable_arrays();
    std::vector<DevArray<float>*> fm;
    if (V.mixed) {
        fm = B->f.mutable_arrays();
        dm.push_back(&B->HA);   // anchor products, rewritten at every FP32 restart
        dm.push_back(&B->GA);
    }
    // Snapshot.
    std::vector<std::unique_ptr<DevArray<double>>> sd;
    std::vector<std::unique_ptr<DevArray<float>>> sf;
    auto copy = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, B->s));
    };
    for (auto* a : dm) {
        sd.emplace_back(new DevArray<double>());
        sd.back()->alloc(a->n);
        copy(sd.back()->p, a->p, a->n * sizeof(double));
    }
    for (auto* a : fm) {
        sf.emplace_back(new DevArray<float>());
        sf.back()->alloc(a->n);
        copy(sf.back()->p, a->p, a->n * sizeof(float));
    }
    DevArray<PdlpState> st0;
    DevArray<Ctl> ctl0;
    DevArray<int> list0, cnt0;
    st0.alloc(B->count);
    ctl0.alloc(B->count);
    list0.alloc(B->count);
    cnt0.alloc(1);
    copy(st0.p, B->st.p, B->count * sizeof(PdlpState));
    copy(ctl0.p, B->ctl.p, B->count * sizeof(Ctl));
    copy(list0.p, B->list.p, B->count * sizeof(int));
    copy(cnt0.p, B->count_d.p, sizeof(int));
    auto restore = [&]() {
        for (size_t k = 0; k < dm.size(); ++k) copy(dm[k]->p, sd[k]->p, dm[k]->n * sizeof(double));
        for (size_t k = 0; k < fm.size(); ++k) copy(fm[k]->p, sf[k]->p, fm[k]->n * sizeof(float));
        copy(B->st.p, st0.p, B->count * sizeof(PdlpState));
        copy(B->ctl.p, ctl0.p, B->count * sizeof(Ctl));
        copy(B->list.p, list0.p, B->count * sizeof(int));
        copy(B->count_d.p, cnt0.p, sizeof(int));
        const int zero = 0;
        HIP_CHECK(hipMemcpyAsync(B->head.p, &zero, sizeof(int), hipMemcpyHostToDevice, B->s));
        HIP_CHECK(hipStreamSynchronize(B->s));
    };
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    int best = 512;
    float best_ms = 1e30f;
    const int saved_wg = B->workgroups;
    for (int t : cands) {
        B->workgroups = std::max(1, std::min(B->count, cu * (2048 / t)));
        HIP_CHECK(hipEventRecord(e0, B->s));
        launch_run(B, V, t);
        HIP_CHECK(hipEventRecord(e1, B->s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
        restore();
    }
    B->workgroups = saved_wg;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

void build_graph(nir_batch* B, int rounds) {
    const View V = B->view();
    HIP_CHECK(hipStreamBeginCapture(B->s, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < rounds; ++r) launch_run(B, V, B->threads);
    HIP_CHECK(hipStreamEndCapture(B->s, &B->graph));
    HIP_CHECK(hipGraphInstantiate(&B->gexec, B->graph, nullptr, nullptr, 0));
}

}  // namespace

extern "C" {

NIR_API void nir_batch_opts_default(nir_batch_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
}

NIR_API int nir_batch_create(nir_ctx* ctx, const nir_lp* const* lps, int32_t count, const nir_params* p,
                             const nir_batch_opts* o, nir_batch** out) {
    return nir_batch_create_warm(ctx, lps, count, nullptr, nullptr, p, o, out);
}

NIR_API int nir_batch_create_warm(nir_ctx* ctx, const nir_lp* const* lps, int32_t count,
                                  const double* const* x0s, const double* const* y0s, const nir_params* p,
                                  const nir_batch_opts* o, nir_batch** out) {
    const nir::TraceRange trace("nir_batch_create");
    if (!ctx || !out) return -1;
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto B = std::make_unique<nir_batch>();
    try {
        if (count <= 0 || !lps) throw std::invalid_argument("empty batch");
        if (!p) throw std::invalid_argument("batch needs parameters");
        B->prm = *p;
        if (o) B->opts = *o;
        if (B->prm.mode != NIR_MODE_PDLP) throw std::invalid_argument("batched solves run pdlp mode only");
        if (B->prm.x0 || B->prm.y0 || B->prm.unscale_row || B->prm.unscale_col)
            throw std::invalid_argument("batched solves take warm starts per instance and no unscaling vectors");
        if (B->opts.threads != 0 && B->opts.threads != 256 && B->opts.threads != 512 && B->opts.threads != 1024)
            throw std::invalid_argument("batch threads must be 0, 256, 512 or 1024");
        if (B->opts.quantum < 0 || B->opts.rounds_per_graph < 0 || B->opts.workgroups < 0)
            throw std::invalid_argument("negative batch option");
        HIP_CHECK(hipSetDevice(ctx->device));
        B->ctx = ctx;
        B->s = ctx->stream;
        B->count = count;
        const bool mixed = B->prm.precision == NIR_PREC_MIXED;

        // Host layout: offsets, CSR(A) by counting transpose, CSR(A^T) = CSC.
        B->h_inst.resize(count);
        B->h_cost.resize(count);
        B->h_order.resize(count);
        B->h_corder.resize(count);
        B->h_offset.resize(count);
        long long xo = 0, yo = 0, nz = 0, arp = 0, trp = 0;
        for (int i = 0; i < count; ++i) {
            const nir_lp* lp = lps[i];
            check_lp(lp);
            constexpr long long kMaxEntries = 1LL << 28;   // 32-bit byte offsets in the device loops
            if (lp->nrow >= kMaxEntries || lp->ncol >= kMaxEntries ||
                (lp->ncol > 0 && lp->col_start[lp->ncol] >= kMaxEntries))
                throw std::invalid_argument("batch instances are limited to 2^28 rows, columns and nonzeros each");
            // The persistent kernel sums each row on one lane in sequential
            // order, which is pdlp mode's order up to kLongRow nonzeros only.
            if (longest_line(lp) > kLongRow)
                throw std::invalid_argument("batch instances with a row or column of more than 4096 nonzeros are not "
                                            "supported: solve them with the single-LP engine");
            Inst& I = B->h_inst[i];
            I.m = lp->nrow;
            I.n = lp->ncol;
            I.nnz = lp->ncol > 0 ? static_cast<int>(lp->col_start[lp->ncol]) : 0;
            I.pad = 0;
            I.arp = arp;
            I.trp = trp;
            I.nz = nz;
            I.xo = xo;
            I.yo = yo;
            arp += I.m + 1;
            trp += I.n + 1;
            nz += I.nnz;
            xo += I.n;
            yo += I.m;
            B->h_cost[i].assign(lp->cost, lp->cost + lp->ncol);
            B->h_offset[i] = lp->obj_offset;
            B->total_nnz += I.nnz;
        }
        B->total_x = xo;
        B->total_y = yo;
        B->total_nz = nz;
        std::vector<int> arow(arp), acol(nz), trow(trp), tcol(nz);
        std::vector<double> aval(nz), tval(nz), cost(xo), collo(xo), colup(xo), rowlo(yo), rowup(yo);
        std::vector<int> colmap(xo);
        {
            // Instances write disjoint ranges: built by host threads, instance
            // i on thread i mod nt.
            const int nt = std::max(1, std::min(16, static_cast<int>(std::thread::hardware_concurrency())));
            std::vector<std::exception_ptr> err(static_cast<size_t>(nt));
            auto build = [&](int t0) {
                try {
                    std::vector<int> ptr, col, pos, cpos;
                    std::vector<double> val;
                    for (int i = t0; i < count; i += nt) {
                        const nir_lp* lp = lps[i];
                        const Inst& I = B->h_inst[i];
                        csr_from_csc(lp, ptr, col, val);
                        // Rows and columns by decreasing length (stable orders),
                        // every row and column keeping its entries in the
                        // caller's order under the new labels.
                        auto by_length = [](std::vector<int>& ord, int k, auto len) {
                            ord.resize(static_cast<size_t>(k));
                            std::iota(ord.begin(), ord.end(), 0);
                            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return len(a) > len(b); });
                        };
                        std::vector<int>& order = B->h_order[i];
                        std::vector<int>& corder = B->h_corder[i];
                        by_length(order, I.m, [&](int r) { return ptr[r + 1] - ptr[r]; });
                        by_length(corder, I.n, [&](int j) {
                            return static_cast<int>(lp->col_start[j + 1] - lp->col_start[j]);
                        });
                        pos.resize(static_cast<size_t>(I.m));
                        cpos.resize(static_cast<size_t>(I.n));
                        for (int r = 0; r < I.m; ++r) pos[order[r]] = r;
                        for (int j = 0; j < I.n; ++j) cpos[corder[j]] = j;
                        int at = 0;
                        arow[I.arp] = 0;
                        for (int r = 0; r < I.m; ++r) {
                            const int old = order[r];
                            for (int p = ptr[old]; p < ptr[old + 1]; ++p, ++at) {
                                acol[I.nz + at] = cpos[col[p]];
                                aval[I.nz + at] = val[p];
                            }
                            arow[I.arp + r + 1] = at;
                            rowlo[I.yo + r] = lp->rowlo[old];
                            rowup[I.yo + r] = lp->rowup[old];
                        }
                        at = 0;
                        trow[I.trp] = 0;
                        for (int j = 0; j < I.n; ++j) {
                            const int old = corder[j];
                            for (long long p = lp->col_start[old]; p < lp->col_start[old + 1]; ++p, ++at) {
                                tcol[I.nz + at] = pos[lp->row_index[p]];
                                tval[I.nz + at] = lp->value[p];
                            }
                            trow[I.trp + j + 1] = at;
                            cost[I.xo + j] = lp->cost[old];
                            collo[I.xo + j] = lp->collo[old];
                            colup[I.xo + j] = lp->colup[old];
                            colmap[I.xo + j] = old;
                        }
                    }
                } catch (...) {
                    err[static_cast<size_t>(t0)] = std::current_exception();
                }
            };
            std::vector<std::thread> pool;
            for (int t = 1; t < nt; ++t) pool.emplace_back(build, t);
            build(0);
            for (auto& t : pool) t.join();
            for (auto& e : err)
                if (e) std::rethrow_exception(e);
        }
        hipStream_t s = B->s;
        auto up_i = [&](DevArray<int>& a, const std::vector<int>& v) { a.alloc(v.size()); a.upload(v.data(), v.size(), s); };
        auto up_d = [&](DevArray<double>& a, const std::vector<double>& v) {
            a.alloc(v.size());
            a.upload(v.data(), v.size(), s);
        };
        B->inst.alloc(count);
        B->inst.upload(B->h_inst.data(), count, s);
        up_i(B->arow, arow);
        up_i(B->acol, acol);
        up_i(B->trow, trow);
        up_i(B->tcol, tcol);
        up_d(B->cost, cost);
        up_d(B->collo, collo);
        up_d(B->colup, colup);
        up_d(B->rowlo, rowlo);
        up_d(B->rowup, rowup);
        up_d(B->offset, B->h_offset);
        up_i(B->colmap, colmap);
        if (x0s || y0s) {
            // Warm starts, concatenated like the column / row vectors.
            B->warm = true;
            std::vector<double> hx(xo, 0.0), hy(yo, 0.0);
            std::vector<int> fx(count, 0), fy(count, 0);
            for (int i = 0; i < count; ++i) {
                const Inst& I = B->h_inst[i];
                if (x0s && x0s[i]) {
                    const std::vector<int>& corder = B->h_corder[i];
                    for (int j = 0; j < I.n; ++j) hx[I.xo + j] = x0s[i][corder[j]];
                    fx[i] = 1;
                }
                if (y0s && y0s[i]) {
                    const std::vector<int>& order = B->h_order[i];
                    for (int r = 0; r < I.m; ++r) hy[I.yo + r] = y0s[i][order[r]];
                    fy[i] = 1;
                }
            }
            up_d(B->x0, hx);
            up_d(B->y0, hy);
            up_i(B->has_x0, fx);
            up_i(B->has_y0, fy);
        }
        B->d.alloc(xo, yo, nz, true);
        B->d.aval.upload(aval.data(), nz, s);
        B->d.tval.upload(tval.data(), nz, s);
        if (mixed) B->f.alloc(xo, yo, nz, false);
        // Sliced indices (written by the setup kernel) in 16 bits when every
        // instance's rows and columns fit.
        {
            int most = 0;
            for (const Inst& I : B->h_inst) most = std::max(most, std::max(I.m, I.n));
            B->idx16 = most <= 65536 ? 1 : 0;
            B->sacol.alloc(static_cast<size_t>(nz) * (B->idx16 ? 2 : 4));
            B->stcol.alloc(static_cast<size_t>(nz) * (B->idx16 ? 2 : 4));
        }
        B->R.alloc(yo);
        B->C.alloc(xo);
        B->hs.alloc(yo);
        if (mixed) {
            B->HA.alloc(yo);
            B->GA.alloc(xo);
        }
        B->st.alloc(count);
        B->ctl.alloc(count);
        // Work list: largest instances first.
        std::vector<int> order(count);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            const Inst& A = B->h_inst[a];
            const Inst& Bb = B->h_inst[b];
            return static_cast<long long>(A.nnz) + A.m + A.n > static_cast<long long>(Bb.nnz) + Bb.m + Bb.n;
        });
        up_i(B->list, order);
        up_i(B->count_d, std::vector<int>{count});
        up_i(B->head, std::vector<int>{0});
        B->launches.alloc(1);
        {
            const long long zero = 0;
            B->launches.upload(&zero, 1, s);
        }
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&B->h_poll), 2 * sizeof(int), hipHostMallocDefault));
        B->h_poll[0] = B->h_poll[1] = count;
        for (auto& e : B->poll_ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));

        const int cu = compute_units(ctx->device);
        // Setup on 512-lane workgroups, then the compaction of the list.
        B->workgroups = B->opts.workgroups > 0 ? B->opts.workgroups : std::max(1, std::min(count, 4 * cu));
        View V = B->view();
        launch_setup(B.get(), V, 512);
        k_batch_compact<<<1, 64, 0, s>>>(V);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(s));
        // LDS for the staged operands of the 1024-lane kernel, when the
        // largest max(m, n) in FP64 fits beside the workgroup's static
        // scratch; otherwise no staging and every instance gathers from
        // global memory.
        {
            constexpr long long kStageCap = 152 * 1024;
            long long need = 0, both = 0;
            for (const Inst& I : B->h_inst) {
                need = std::max<long long>(need, std::max(I.m, I.n));
                both = std::max<long long>(both, static_cast<long long>(I.m) + I.n);
            }
            // One region of max(m, n) must fit every instance; instances
            // whose m + n entries also fit keep y and xbar apart.
            B->stage_bytes =
                need * 8 <= kStageCap ? static_cast<int>(std::max<long long>(std::min(both * 8, kStageCap), 8)) : 0;
            if (B->stage_bytes > 0 &&
                hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_run<1024, true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, B->stage_bytes) != hipSuccess) {
                (void)hipGetLastError();
                B->stage_bytes = 0;
            }
        }
        B->threads = B->opts.threads > 0 ? B->opts.threads : autotune_threads(B.get(), cu);
        if (B->opts.workgroups <= 0) B->workgroups = std::max(1, std::min(count, cu * (2048 / B->threads)));
        {
            const long long zero = 0;
            B->launches.upload(&zero, 1, s);
        }
        HIP_CHECK(hipStreamSynchronize(s));
        B->setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = B.release();
        return 0;
    } catch (const std::exception& ex) {
        ctx->error = ex.what();
        // Setup kernels and uploads may still be queued against B's buffers.
        if (B->s) (void)hipStreamSynchronize(B->s);
        return -1;
    }
}

NIR_API int nir_batch_solve(nir_batch* B) {
    const nir::TraceRange trace("nir_batch_solve");
    if (!B) return -1;
    try {
        HIP_CHECK(hipSetDevice(B->ctx->device));
        auto t0 = std::chrono::steady_clock::now();
        const int rounds = B->opts.rounds_per_graph > 0 ? B->opts.rounds_per_graph : 4;
        const View V = B->view();
        long long serial = 0;
        int pending[2] = {0, 0};
        HIP_CHECK(hipMemcpy(&B->h_poll[0], B->count_d.p, sizeof(int), hipMemcpyDeviceToHost));
        int remaining = B->h_poll[0];
        // Every launch runs `quantum` blocks of each listed instance and the
        // iteration cap bounds every instance, so the list must be empty
        // after this many replays (two more for the one in flight).
        const View V0 = B->view();
        const long long launches_bound =
            (V0.max_iters + static_cast<long long>(V0.eval_freq) * V0.quantum - 1) /
                (static_cast<long long>(V0.eval_freq) * V0.quantum) + 2;
        const long long replay_bound = (launches_bound + rounds - 1) / rounds + 2;
        long long replays = 0;
        while (remaining > 0) {
            if (++replays > replay_bound) throw std::runtime_error("batch did not drain within its iteration bound");
            if (B->prm.should_stop && B->prm.should_stop(B->prm.stop_user)) {
                B->stopped = true;
                break;
            }
            if (B->prm.use_graph) {
                if (!B->gexec) build_graph(B, rounds);
                HIP_CHECK(hipGraphLaunch(B->gexec, B->s));
            } else {
                for (int r = 0; r < rounds; ++r) launch_run(B, V, B->threads);
            }
            const int slot = static_cast<int>(serial++ & 1);
            HIP_CHECK(hipMemcpyAsync(&B->h_poll[slot], B->count_d.p, sizeof(int), hipMemcpyDeviceToHost, B->s));
            HIP_CHECK(hipEventRecord(B->poll_ev[slot], B->s));
            pending[slot] = 1;
            // Keep one replay in flight: wait for the previous one.
            const int prev = slot ^ 1;
            if (pending[prev]) {
                HIP_CHECK(hipEventSynchronize(B->poll_ev[prev]));
                pending[prev] = 0;
                remaining = B->h_poll[prev];
                if (B->prm.verbose > 1)
                    std::fprintf(stderr, "[batch] %.4f s: %d instances unfinished\n",
                                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(),
                                 remaining);
            }
        }
        HIP_CHECK(hipStreamSynchronize(B->s));
        B->solved = true;
        B->solve_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    } catch (const std::exception& ex) {
        B->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_batch_result(nir_batch* B, int32_t i, double* x, double* y, nir_info* info) {
    if (!B || i < 0 || i >= B->count) return -1;
    try {
        HIP_CHECK(hipSetDevice(B->ctx->device));
        const Inst& I = B->h_inst[i];
        PdlpState h{};
        Ctl c{};
        HIP_CHECK(hipMemcpy(&h, B->st.p + i, sizeof(h), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(&c, B->ctl.p + i, sizeof(c), hipMemcpyDeviceToHost));
        std::vector<double> xs(static_cast<size_t>(I.n)), ys(static_cast<size_t>(I.m));
        if (c.fp32) {
            // T(z) of the FP32 phase: FP64 anchor + FP32 delta.
            std::vector<float> xf(static_cast<size_t>(I.n)), yf(static_cast<size_t>(I.m));
            if (I.n) HIP_CHECK(hipMemcpy(xf.data(), B->f.xp.p + I.xo, I.n * sizeof(float), hipMemcpyDeviceToHost));
            if (I.m) HIP_CHECK(hipMemcpy(yf.data(), B->f.yp.p + I.yo, I.m * sizeof(float), hipMemcpyDeviceToHost));
            if (I.n) HIP_CHECK(hipMemcpy(xs.data(), B->d.xa.p + I.xo, I.n * sizeof(double), hipMemcpyDeviceToHost));
            if (I.m) HIP_CHECK(hipMemcpy(ys.data(), B->d.ya.p + I.yo, I.m * sizeof(double), hipMemcpyDeviceToHost));
            for (int j = 0; j < I.n; ++j) xs[j] = xs[j] + static_cast<double>(xf[j]);
            for (int r = 0; r < I.m; ++r) ys[r] = ys[r] + static_cast<double>(yf[r]);
        } else {
            if (I.n) HIP_CHECK(hipMemcpy(xs.data(), B->d.xp.p + I.xo, I.n * sizeof(double), hipMemcpyDeviceToHost));
            if (I.m) HIP_CHECK(hipMemcpy(ys.data(), B->d.yp.p + I.yo, I.m * sizeof(double), hipMemcpyDeviceToHost));
        }
        std::vector<double> R(static_cast<size_t>(I.m)), Cc(static_cast<size_t>(I.n));
        if (I.m) HIP_CHECK(hipMemcpy(R.data(), B->R.p + I.yo, I.m * sizeof(double), hipMemcpyDeviceToHost));
        if (I.n) HIP_CHECK(hipMemcpy(Cc.data(), B->C.p + I.xo, I.n * sizeof(double), hipMemcpyDeviceToHost));
        {
            // Device column j is the caller's column corder[j].
            const std::vector<int>& corder = B->h_corder[i];
            std::vector<double> xo(static_cast<size_t>(I.n));
            for (int j = 0; j < I.n; ++j) xo[corder[j]] = xs[j] * Cc[j];
            xs.swap(xo);
        }
        if (x) std::copy(xs.begin(), xs.end(), x);
        if (y) {
            const std::vector<int>& order = B->h_order[i];   // device row r is the caller's row order[r]
            for (int r = 0; r < I.m; ++r) y[order[r]] = ys[r] * R[r];
        }
        if (info) {
            std::memset(info, 0, sizeof(*info));
            PdlpState fin = h;
            fin.status = c.status;
            info->status = c.finished ? sov_status(fin) : (B->stopped ? 4 : 6);
            info->iterations = h.total;
            info->polish_iterations = c.fp64_at >= 0 ? h.total - c.fp64_at : 0;
            info->restarts = h.restarts;
            DDHost acc(B->h_offset[i]);
            const std::vector<double>& cost = B->h_cost[i];
            for (int j = 0; j < I.n; ++j)
                if (cost[j] != 0.0 && xs[j] != 0.0) acc.add_prod(cost[j], xs[j]);
            info->objective = acc.value();
            info->tau = h.tau;
            info->sigma = h.sigma;
            info->primal_weight = h.omega;
            info->rel_primal = h.rel_p;
            info->rel_dual = h.rel_d;
            info->rel_gap = h.rel_g;
            info->setup_seconds = B->setup_s;
            info->solve_seconds = B->solve_s;
        }
        return 0;
    } catch (const std::exception& ex) {
        B->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_batch_get_stats(nir_batch* B, nir_batch_stats* out) {
    if (!B || !out) return -1;
    try {
        HIP_CHECK(hipSetDevice(B->ctx->device));
        std::memset(out, 0, sizeof(*out));
        out->count = B->count;
        out->threads = B->threads;
        out->workgroups = B->workgroups;
        out->quantum = B->opts.quantum > 0 ? B->opts.quantum : kDefaultQuantum;
        long long launches = 0;
        HIP_CHECK(hipMemcpy(&launches, B->launches.p, sizeof(launches), hipMemcpyDeviceToHost));
        out->launches = launches;
        std::vector<PdlpState> h(static_cast<size_t>(B->count));
        HIP_CHECK(hipMemcpy(h.data(), B->st.p, B->count * sizeof(PdlpState), hipMemcpyDeviceToHost));
        double nzit = 0.0;
        long long it = 0;
        for (int i = 0; i < B->count; ++i) {
            it += h[i].total;
            nzit += static_cast<double>(h[i].total) * B->h_inst[i].nnz;
        }
        out->total_iterations = it;
        out->nnz_iterations = nzit;
        out->total_nnz = B->total_nnz;
        out->setup_seconds = B->setup_s;
        out->solve_seconds = B->solve_s;
        return 0;
    } catch (const std::exception& ex) {
        B->ctx->error = ex.what();
        return -1;
    }
}

NIR_API void nir_batch_destroy(nir_batch* B) {
    if (!B) return;
    (void)hipSetDevice(B->ctx->device);
    // A solve that returned an error may leave a graph replay and the poll
    // copy into the pinned word in flight; they use the buffers, graph and
    // events the destructor frees.
    (void)hipStreamSynchronize(B->s);
    delete B;
}

}  // extern "C"
