// PDHG engine: device-resident LP sessions on one MI355X.
//
// Reference mode restates src/gpu/pdhg_solver.hpp:25-107 on the device:
//   power iteration for ||A||_2 (12 rounds, :32-45), x at the bound midpoint
//   (:47-57), then per iteration
//     primal kernel : g = A^T y over CSR(A^T) (the CSC arrays of A), fused
//                     x+ = clamp(x - tau (c - g)), xbar = 2 x+ - x, x = x+
//                     (:70-78)
//     dual kernel   : h = A xbar over CSR(A), fused
//                     y += sigma (clamp(h - y/sigma) - h) and the max row
//                     violation of h (:82-95)
//   and the reference's stop rule (k > 20, k % 32 == 0, violation <= 1e-6,
//   :97-100) decided on the device: the next primal kernel reads the
//   violation slot and raises a done flag that every later kernel honours.
// The host only enqueues iteration blocks (hipGraph-captured) and polls the
// done flag once per 256 iterations, which is also where the deadline is
// checked (:63). Each product may run as several column-segment passes
// (spmv.hpp); the fused epilogue runs in the last pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../../include/nirnaya/pdhg.h"
#include "devmat.hpp"
#include "ctx.hpp"
#include "lpcheck.hpp"
#include "pdlp.hpp"
#include "trace.hpp"
#include "verify.hpp"

namespace nir {
namespace gpu {

// ------------------------------------------------------------------ kernels

struct RefState {
    int done;      // raised when the stop rule fires
    int base;      // iteration index of k_lo == 0
    int iters;     // iterations completed when done was raised
    int pad;
    unsigned long long vslot[2];   // max row violation of iteration k in slot k & 1 (bits of a double >= 0)
};

// y = M v. ORDER and LONG as in stream_rows (ORDER 1: pdlp mode's order
// with split rows; LONG 0: the pdlp kernels' compilation for layouts without
// long-row blocks), so that a pdlp layout candidate is timed with the code
// its iterations will run.
template <int TILE, typename MV, typename T, int ORDER = 0, int LONG = 2>
__global__ __launch_bounds__(kThreads) void k_spmv_store(MV P, const T* __restrict__ v, T* __restrict__ out) {
    __shared__ T prod[lds_slots(TILE)];
    if constexpr (ORDER == 1) stream_rows<TILE, 1, LONG>(P, v, prod, blockIdx.x, [&](int r, T s) { out[r] = s; });
    else stream_rows<TILE>(P, v, prod, blockIdx.x, [&](int r, T s) { out[r] = s; });
}

// Segment passes (MV = Pass) are compiled twice, as in pdlp mode: LONG = 0
// without the long-row code for layouts that have no block of one row longer
// than a tile, LONG = 2 with it otherwise. Fused-segment launches (SegCsr)
// hold no long rows.
template <int TILE, typename MV, int LONG = 2>
__global__ __launch_bounds__(kThreads) void k_primal_ref(MV P, const double* __restrict__ y,
                                                         double* __restrict__ x, double* __restrict__ xbar,
                                                         const double* __restrict__ c,
                                                         const double* __restrict__ lo,
                                                         const double* __restrict__ up, double tau,
                                                         RefState* st, int k_lo, int check) {
    __shared__ double prod[lds_slots(TILE)];
    if (st->done) return;
    const int k = st->base + k_lo;
    if (P.first) {
        // The reference tests its stop rule at the end of iteration k-1.
        if (check && k > 21 && ((k - 1) & 31) == 0) {
            const double V = __longlong_as_double(static_cast<long long>(st->vslot[(k - 1) & 1]));
            if (V <= 1e-6) {
                if (blockIdx.x == 0 && threadIdx.x == 0) {
                    st->done = 1;
                    st->iters = k;
                }
                return;
            }
        }
    }
    struct Ops {
        double x, c, l, u;
    };
    const auto pre = [&](int j) {
        Ops o{0.0, 0.0, 0.0, 0.0};
        if (P.last) o = Ops{x[j], c[j], lo[j], up[j]};
        return o;
    };
    const auto on_row = [&](int j, double g, const Ops& o) {
        const double xj = o.x;
        double cand = xj - tau * (o.c - g);
        if (finite_bound(o.l)) cand = amax(o.l, cand);
        if (finite_bound(o.u)) cand = amin(o.u, cand);
        xbar[j] = 2.0 * cand - xj;
        x[j] = cand;
    };
    if constexpr (std::is_same<MV, Pass<double>>::value)
        stream_rows<TILE, 0, LONG, LONG == 2 ? 16 : 8>(P, y, prod, blockIdx.x, pre, on_row);
    else stream_rows<TILE>(P, y, prod, blockIdx.x, pre, on_row);
}

template <int TILE, typename MV, int LONG = 2>
__global__ __launch_bounds__(kThreads) void k_dual_ref(MV P, const double* __restrict__ xbar,
                                                       double* __restrict__ y, const double* __restrict__ rl,
                                                       const double* __restrict__ ru, double sigma,
                                                       const RefState* st, double* __restrict__ vpart) {
    __shared__ double prod[lds_slots(TILE)];
    __shared__ double red[kWaves];
    if (st->done) return;
    double vmax = 0.0;
    struct Ops {
        double y, l, u;
    };
    const auto pre = [&](int i) {
        Ops o{0.0, 0.0, 0.0};
        if (P.last) o = Ops{y[i], rl[i], ru[i]};
        return o;
    };
    const auto on_row = [&](int i, double h, const Ops& o) {
        const double yi = o.y;
        double r = h - yi / sigma;
        const double l = o.l, u = o.u;
        const bool lf = finite_bound(l), uf = finite_bound(u);
        if (lf) r = amax(l, r);
        if (uf) r = amin(u, r);
        y[i] = yi + sigma * (r - h);
        double viol = 0.0;
        if (lf) viol = amax(viol, l - h);
        if (uf) viol = amax(viol, h - u);
        vmax = amax(vmax, viol);
    };
    if constexpr (std::is_same<MV, Pass<double>>::value)
        stream_rows<TILE, 0, LONG, LONG == 2 ? 16 : 8>(P, xbar, prod, blockIdx.x, pre, on_row);
    else stream_rows<TILE>(P, xbar, prod, blockIdx.x, pre, on_row);
    if (P.last) {
        // One plain store per block; the maximum over blocks is only formed on
        // the iterations the stop rule reads it (k_vmax_reduce).
        vmax = block_max(vmax, red);
        if (threadIdx.x == 0) vpart[blockIdx.x] = vmax;
    }
}

// Max row violation of iteration k = base + k_lo over the dual kernel's
// per-block maxima, into slot k & 1 (one workgroup, fixed order).
__global__ __launch_bounds__(kThreads) void k_vmax_reduce(const double* __restrict__ vpart, int nb, RefState* st,
                                                          int k_lo) {
    __shared__ double red[kWaves];
    if (st->done) return;
    const int k = st->base + k_lo;
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) v = amax(v, vpart[b]);
    v = block_max(v, red);
    if (threadIdx.x == 0) st->vslot[k & 1] = static_cast<unsigned long long>(__double_as_longlong(v));
}

// sum_{p in [p0, p1)} val[p] v[col[p]] from 0.0 in ascending p, products
// rounded before the add; four operand reads in flight per step.
__device__ __forceinline__ double seq_dot(int p0, int p1, const int* __restrict__ col, const double* __restrict__ val,
                                          const double* __restrict__ v) {
    double s = 0.0;
    int p = p0;
    for (; p + 4 <= p1; p += 4) {
        const int c0 = col[p], c1 = col[p + 1], c2 = col[p + 2], c3 = col[p + 3];
        const double a0 = val[p], a1 = val[p + 1], a2 = val[p + 2], a3 = val[p + 3];
        const double g0 = v[c0], g1 = v[c1], g2 = v[c2], g3 = v[c3];
        s += a0 * g0;
        s += a1 * g1;
        s += a2 * g2;
        s += a3 * g3;
    }
    for (; p < p1; ++p) s += val[p] * v[col[p]];
    return s;
}

// Small LPs: `count` reference iterations k0, k0 + 1, ... in one workgroup,
// with x, xbar and y resident in LDS (2n + m doubles) and a barrier between
// the primal and the dual half instead of a kernel boundary. Lanes own
// columns and rows by stride; every product is the sequential sum of the
// multi-kernel path (ascending row order for A^T y, ascending column order
// for A xbar, from 0.0), and the updates are the same expressions, so the
// iterates are the same bits. The stop rule is tested after the dual half of
// iteration k (the reference's `iter > 20 && iter % 32 == 0 && V <= 1e-6`,
// pdhg_solver.hpp:97-100); vpart[0] receives the last iteration's maximum
// row violation (the other entries hold 0).
template <int MAT>
__global__ __launch_bounds__(1024) void k_ref_small(int m, int n, const int* __restrict__ trp,
                                                    const int* __restrict__ tcol, const double* __restrict__ tval,
                                                    const int* __restrict__ arp, const int* __restrict__ acol,
                                                    const double* __restrict__ aval, double* __restrict__ x,
                                                    double* __restrict__ xbar, double* __restrict__ y,
                                                    const double* __restrict__ c, const double* __restrict__ lo,
                                                    const double* __restrict__ up, const double* __restrict__ rl,
                                                    const double* __restrict__ ru, double tau, double sigma,
                                                    RefState* st, double* __restrict__ vpart, int nvpart, int k0,
                                                    int count, int check) {
    extern __shared__ double lds[];
    __shared__ double red[16];
    __shared__ int stop;
    if (st->done) return;
    double* sx = lds;
    double* sxb = lds + n;
    double* sy = lds + 2 * static_cast<size_t>(n);
    const int tid = threadIdx.x, nt = blockDim.x, nw = nt >> 6;
    for (int j = tid; j < n; j += nt) {
        sx[j] = x[j];
        sxb[j] = xbar[j];
    }
    for (int i = tid; i < m; i += nt) sy[i] = y[i];
    const int nz = MAT ? arp[m] : 0;   // from global memory, before arp moves to LDS
    if (MAT) {
        // MAT >= 1: both matrices copied into LDS behind the vectors (values,
        // then row pointers and indices); MAT == 2: the costs and bounds as
        // well, so nothing in the iteration loop waits on global memory.
        double* sav = sy + m;
        double* stv = sav + nz;
        int* sarp = reinterpret_cast<int*>(stv + nz);
        int* strp = sarp + (m + 1);
        int* sac = strp + (n + 1);
        int* stc = sac + nz;
        for (int p = tid; p < nz; p += nt) {
            sav[p] = aval[p];
            stv[p] = tval[p];
            sac[p] = acol[p];
            stc[p] = tcol[p];
        }
        for (int i = tid; i <= m; i += nt) sarp[i] = arp[i];
        for (int j = tid; j <= n; j += nt) strp[j] = trp[j];
        aval = sav;
        tval = stv;
        arp = sarp;
        trp = strp;
        acol = sac;
        tcol = stc;
    }
    if (MAT == 2) {
        // then the costs and bounds (doubles, after the index arrays)
        const size_t ints = static_cast<size_t>(m) + 1 + n + 1 + 2 * static_cast<size_t>(nz);
        double* sconst = reinterpret_cast<double*>(const_cast<int*>(arp)) + (ints + 1) / 2;
        double *sc = sconst, *slo = sc + n, *sup = slo + n, *srl = sup + n, *sru = srl + m;
        for (int j = tid; j < n; j += nt) {
            sc[j] = c[j];
            slo[j] = lo[j];
            sup[j] = up[j];
        }
        for (int i = tid; i < m; i += nt) {
            srl[i] = rl[i];
            sru[i] = ru[i];
        }
        c = sc;
        lo = slo;
        up = sup;
        rl = srl;
        ru = sru;
    }
    if (tid == 0) stop = 0;
    __syncthreads();
    double V = 0.0;
    int k = k0;
    for (; k < k0 + count; ++k) {
        for (int j = tid; j < n; j += nt) {
            const double g = seq_dot(trp[j], trp[j + 1], tcol, tval, sy);
            const double xj = sx[j];
            double cand = xj - tau * (c[j] - g);
            if (finite_bound(lo[j])) cand = amax(lo[j], cand);
            if (finite_bound(up[j])) cand = amin(up[j], cand);
            sxb[j] = 2.0 * cand - xj;
            sx[j] = cand;
        }
        __syncthreads();
        double vloc = 0.0;
        for (int i = tid; i < m; i += nt) {
            const double h = seq_dot(arp[i], arp[i + 1], acol, aval, sxb);
            const double yi = sy[i];
            double r = h - yi / sigma;
            const double l = rl[i], u = ru[i];
            const bool lf = finite_bound(l), uf = finite_bound(u);
            if (lf) r = amax(l, r);
            if (uf) r = amin(u, r);
            sy[i] = yi + sigma * (r - h);
            double viol = 0.0;
            if (lf) viol = amax(viol, l - h);
            if (uf) viol = amax(viol, h - u);
            vloc = amax(vloc, viol);
        }
        // Maximum violation over the workgroup (every lane gets it); the
        // barriers also order this dual half before the next primal half.
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) vloc = amax(vloc, __shfl_xor(vloc, o));
        if ((tid & 63) == 0) red[tid >> 6] = vloc;
        __syncthreads();
        V = red[0];
        for (int w = 1; w < nw; ++w) V = amax(V, red[w]);
        __syncthreads();
        // The violation slot the tiled path's stop rule reads (k_vmax_reduce),
        // kept current so that a session may continue on either path.
        if ((k & 31) == 0 && tid == 0) st->vslot[k & 1] = static_cast<unsigned long long>(__double_as_longlong(V));
        if (check && k > 20 && (k & 31) == 0 && V <= 1e-6) {
            ++k;
            if (tid == 0) stop = 1;
            break;
        }
    }
    for (int j = tid; j < n; j += nt) {
        x[j] = sx[j];
        xbar[j] = sxb[j];
    }
    for (int i = tid; i < m; i += nt) y[i] = sy[i];
    for (int b = 1 + tid; b < nvpart; b += nt) vpart[b] = 0.0;
    if (tid == 0) {
        vpart[0] = V;
        if (stop) {
            st->done = 1;
            st->iters = k;
        }
    }
}

__global__ void k_set_base(RefState* st, int base) { st->base = base; }
__global__ void k_advance(RefState* st, int by) { st->base += by; }

// Plain grid-stride copy: the calibration kernel for HBM byte counters.
__global__ __launch_bounds__(256) void k_copy_calibrate(const double4* __restrict__ in, double4* __restrict__ out,
                                                        long long n) {
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = in[i];
}

// Access floor of a product: every stored nonzero of slots [begin, end) reads
// its value and column index and gathers v[col] once, with no row structure,
// no LDS and no epilogue - the memory accesses any CSR product over this
// layout must make, at the rate the chip sustains for them. Per-lane sums go
// to out so the loads stay live.
__global__ __launch_bounds__(256) void k_gather_floor(const int* __restrict__ col, const double* __restrict__ val,
                                                      const double* __restrict__ v, long long begin, long long end,
                                                      double* __restrict__ out) {
    double s = 0.0;
    const long long stride = static_cast<long long>(gridDim.x) * 256;
    for (long long p = begin + static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; p < end; p += stride)
        s += load_stream(val + p) * v[load_stream(col + p)];
    out[static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x] = s;
}

// Random-gather rate probe: n uniform random indices into an operand of len
// entries (splitmix64 of the position), values in [0, 1).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_fill_gather_probe(int* __restrict__ idx, double* __restrict__ val,
                                                           double* __restrict__ x, long long n, int len) {
    const long long stride = static_cast<long long>(gridDim.x) * 256;
    for (long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
        const unsigned long long h = mix64(static_cast<unsigned long long>(i));
        idx[i] = static_cast<int>(h % static_cast<unsigned long long>(len));
        val[i] = static_cast<double>(h >> 11) * 0x1.0p-53;
        if (i < len) x[i] = val[i];
    }
}

template <int WITH_VALUES>
__global__ __launch_bounds__(256) void k_gather_probe(const int* __restrict__ idx, const double* __restrict__ val,
                                                      const double* __restrict__ x, long long n,
                                                      double* __restrict__ out) {
    double s = 0.0;
    const long long stride = static_cast<long long>(gridDim.x) * 256;
    for (long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
        if (WITH_VALUES == 2) {   // streams only: the index is used as a number, nothing gathered
            s += load_stream(val + i) + static_cast<double>(load_stream(idx + i));
        } else {
            const double g = x[load_stream(idx + i)];
            s += WITH_VALUES ? load_stream(val + i) * g : g;
        }
    }
    out[static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x] = s;
}

bool host_finite(double b) { return std::isfinite(b) && std::fabs(b) < 1e20; }

// y = M v over a segmented shape with values `val` (and running sums `part`).
template <typename T>
void spmv_store(const CsrShape& S, const T* val, T* part, const T* v, T* out, hipStream_t s) {
    if (S.fused_ok) {
        const SegCsr<T> F = S.fused(val);
        TILE_DISPATCH(S.shape.tile, k_spmv_store<TL><<<F.nblk, kThreads, 0, s>>>(F, v, out));
    } else {
        for (int p = 0; p < S.nseg; ++p) {
            const Pass<T> P = S.pass(val, part, p);
            // A layout built for pdlp mode (split rows) is timed in its order.
            if (S.shape.split && S.nlong[p] == 0) {
                TILE_DISPATCH(S.shape.tile,
                              (k_spmv_store<TL, Pass<T>, T, 1, 0><<<P.M.nblk, kThreads, 0, s>>>(P, v, out)));
            } else if (S.shape.split) {
                TILE_DISPATCH(S.shape.tile,
                              (k_spmv_store<TL, Pass<T>, T, 1, 2><<<P.M.nblk, kThreads, 0, s>>>(P, v, out)));
            } else if (S.nlong[p] == 0) {
                TILE_DISPATCH(S.shape.tile,
                              (k_spmv_store<TL, Pass<T>, T, 0, 0><<<P.M.nblk, kThreads, 0, s>>>(P, v, out)));
            } else {
                TILE_DISPATCH(S.shape.tile, k_spmv_store<TL><<<P.M.nblk, kThreads, 0, s>>>(P, v, out));
            }
        }
    }
    HIP_CHECK(hipGetLastError());
}

void spmv_store(const DevMatrix& M, const double* v, double* out, hipStream_t s) {
    spmv_store<double>(M.shape, M.val.p, M.part.p, v, out, s);
}

// A segment width's host layout and its values permuted into it, in
// precision T.
template <typename T>
struct WidthLayout {
    SegHost H;
    std::vector<T> vals;
};

template <typename T>
WidthLayout<T> build_width(int nrows, int ncols, const std::vector<int>& rowptr, const int* col, const double* val,
                           int width) {
    WidthLayout<T> L;
    L.H = build_segments(nrows, ncols, rowptr, col, width);
    L.vals.resize(L.H.perm.size());
    for (size_t q = 0; q < L.H.perm.size(); ++q) L.vals[q] = static_cast<T>(val[L.H.perm[q]]);
    std::vector<int>().swap(L.H.perm);   // not needed for timing
    return L;
}

// Time 8 products y = M v for every candidate of one segment width: the
// segmented structure and the permuted values are uploaded once for the
// width, only the row blocks change per tile shape. Keeps the fastest in
// `best` / `best_ms`.
template <typename T>
void time_width(const WidthLayout<T>& L, const std::vector<TileShape>& cands, hipStream_t s, hipEvent_t e0,
                hipEvent_t e1, DevArray<T>& part, DevArray<T>& v, DevArray<T>& out, TileShape& best,
                float& best_ms) {
    CsrShape S;
    upload_structure(S, L.H, s);
    DevArray<T> dval;
    dval.alloc(L.vals.size());
    dval.upload(L.vals.data(), L.vals.size(), s);
    for (const TileShape& c : cands) {
        // The previous candidate's products have completed (event synchronised
        // below), so its block arrays can be replaced.
        set_blocks(S, L.H, c, s);
        for (int w = 0; w < 2; ++w) spmv_store<T>(S, dval.p, part.p, v.p, out.p, s);
        HIP_CHECK(hipEventRecord(e0, s));
        for (int w = 0; w < 8; ++w) spmv_store<T>(S, dval.p, part.p, v.p, out.p, s);
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    HIP_CHECK(hipStreamSynchronize(s));
}

template <typename T>
TileShape time_candidates(int nrows, int ncols, const std::vector<int>& rowptr, const int* col, const double* val,
                          const std::vector<int>& widths, const std::vector<std::vector<TileShape>>& per_width,
                          hipStream_t s) {
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    DevArray<T> part, v, out;
    part.alloc(static_cast<size_t>(std::max(nrows, 1)));
    v.alloc(static_cast<size_t>(std::max(ncols, 1)));
    out.alloc(static_cast<size_t>(std::max(nrows, 1)));
    {
        std::vector<T> ones(static_cast<size_t>(std::max(ncols, 1)), T(1));
        v.upload(ones.data(), ones.size(), s);
    }
    TileShape best = per_width.front().front();
    float best_ms = 1e30f;
    // The host layouts of all widths are built concurrently (O(nnz) each);
    // the products are then timed one width at a time on the stream.
    std::vector<std::future<WidthLayout<T>>> layouts;
    for (int w : widths)
        layouts.push_back(std::async(std::launch::async, [&, w] { return build_width<T>(nrows, ncols, rowptr, col, val, w); }));
    try {
        for (size_t k = 0; k < widths.size(); ++k) {
            const WidthLayout<T> L = layouts[k].get();
            time_width<T>(L, per_width[k], s, e0, e1, part, v, out, best, best_ms);
        }
    } catch (...) {
        for (auto& f : layouts)
            if (f.valid()) f.wait();
        (void)hipStreamSynchronize(s);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

// Autotuner: time y = M v for every candidate layout (segment width, tile,
// fused or per-segment launches) on the real matrix, in the precision the
// iterations will run in (FP32 for mixed precision), and keep the fastest.
// Results are cached per matrix (its shape, a hash of its row pointers and
// of sampled column indices) and precision, so two matrices of one size (A
// and the A^T of a square A, or two LPs solved in one process) are timed
// separately. The choice never changes results: every row, of any length,
// is summed in sequential column order.
unsigned long long layout_signature(int nrows, const std::vector<int>& rowptr, const int* col) {
    unsigned long long h = 0x243f6a8885a308d3ull;
    auto mix = [&](unsigned long long v) {
        h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 33;
    };
    for (int r = 0; r <= nrows; ++r) mix(static_cast<unsigned long long>(static_cast<unsigned>(rowptr[r])));
    const long long nnz = rowptr[nrows];
    const long long step = std::max<long long>(1, nnz / 4096);
    for (long long p = 0; p < nnz; p += step) mix(static_cast<unsigned long long>(static_cast<unsigned>(col[p])));
    return h;
}

TileShape autotune(int nrows, int ncols, const std::vector<int>& rowptr, const int* col, const double* val,
                   hipStream_t s, bool fp32, bool allow_fused, bool allow_segments, bool split) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int, unsigned long long>, TileShape> cache;
    const auto key = std::make_tuple(nrows, ncols, rowptr.back(),
                                     (fp32 ? 1 : 0) + (allow_fused ? 2 : 0) + (allow_segments ? 4 : 0) + (split ? 8 : 0),
                                     layout_signature(nrows, rowptr, col));
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    // Power-of-two segment widths plus the even splits in two, three and
    // four (multiples of 1024 columns): equal segments keep every pass's
    // operand slice the same size (an FP32 operand of a few MB per segment
    // fits one XCD's L2 in two passes).
    std::vector<int> all = {0, 1 << 19, 1 << 18, 1 << 17};
    for (int parts = 2; parts <= 4; ++parts) {
        const int w = static_cast<int>((static_cast<long long>(ncols) / parts + 1023) / 1024 * 1024);
        if (w > 0 && std::find(all.begin(), all.end(), w) == all.end()) all.push_back(w);
    }
    std::vector<int> widths;
    std::vector<std::vector<TileShape>> per_width;
    for (int w : all) {
        if (w != 0 && (w >= ncols || !allow_segments)) continue;
        std::vector<TileShape> cands;
        for (int t : {1024, 2048, 4096})
            for (int f : {0, 1})
                if (f == 0 || (w > 0 && allow_fused)) cands.push_back(TileShape{t, 256, w, f, split ? 1 : 0});
        widths.push_back(w);
        per_width.push_back(std::move(cands));
    }
    const TileShape best = fp32 ? time_candidates<float>(nrows, ncols, rowptr, col, val, widths, per_width, s)
                                : time_candidates<double>(nrows, ncols, rowptr, col, val, widths, per_width, s);
    std::lock_guard<std::mutex> lock(mu);
    cache[key] = best;
    return best;
}

}  // namespace gpu
}  // namespace nir

using namespace nir::gpu;

// ------------------------------------------------------------------ context

static thread_local std::string g_last_error;


struct nir_session {
    nir_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    nir_params prm{};
    int m = 0, n = 0;
    long long nnz = 0;
    double obj_offset = 0.0;
    std::vector<double> h_cost;

    // matrices: CSR(A) for A xbar, CSR(A^T) = CSC(A) for A^T y
    DevMatrix A, AT;
    // vectors
    DevArray<double> c, lo, up, rl, ru, x, xbar, y;
    DevArray<RefState> st;
    DevArray<double> vpart;       // per-block max row violation of the last dual kernel
    RefState* h_poll = nullptr;   // pinned: two poll slots

    // progress
    long long k_next = 0;     // next iteration index to enqueue
    long long max_iters = 0;
    int base_host = 0;
    bool finished = false;
    int status = 0;
    long long iters_done = 0;
    double tau = 0.0, sigma = 0.0;
    double setup_seconds = 0.0, solve_seconds = 0.0;

    // small LPs: all iterations of a chunk in one workgroup (k_ref_small)
    int small_threads = 0;   // 0: the multi-kernel path
    int small_mat = 0;       // 1: the matrices are copied into LDS as well, 2: and costs and bounds
    size_t small_lds = 0;

    // graph of 32 reference iterations
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t poll_ev[2] = {nullptr, nullptr};
    int poll_pending[2] = {0, 0};
    long long poll_serial = 0;

    // PDLP engine state (null in reference mode)
    std::unique_ptr<nir::gpu::Pdlp, nir::gpu::PdlpDeleter> pdlp;

    ~nir_session();
};

nir_session::~nir_session() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& e : poll_ev)
        if (e) (void)hipEventDestroy(e);
    if (h_poll) (void)hipHostFree(h_poll);
    pdlp.reset();
}

namespace {

void launch_primal_ref(nir_session* S, int k_lo) {
    const int check = S->prm.check_termination ? 1 : 0;
    const DevMatrix& M = S->AT;
    if (M.shape.fused_ok) {
        const SegCsr<double> F = M.shape.fused(M.val.p);
        TILE_DISPATCH(M.shape.shape.tile,
                      k_primal_ref<TL><<<F.nblk, kThreads, 0, S->s>>>(F, S->y.p, S->x.p, S->xbar.p, S->c.p, S->lo.p,
                                                                     S->up.p, S->tau, S->st.p, k_lo, check));
        HIP_CHECK(hipGetLastError());
        return;
    }
    for (int p = 0; p < M.shape.nseg; ++p) {
        const Pass<double> P = M.shape.pass(M.val.p, M.part.p, p);
        if (M.shape.nlong[p] > 0) {
            TILE_DISPATCH(M.shape.shape.tile,
                          (k_primal_ref<TL, Pass<double>, 2><<<P.M.nblk, kThreads, 0, S->s>>>(
                              P, S->y.p, S->x.p, S->xbar.p, S->c.p, S->lo.p, S->up.p, S->tau, S->st.p, k_lo, check)));
        } else {
            TILE_DISPATCH(M.shape.shape.tile,
                          (k_primal_ref<TL, Pass<double>, 0><<<P.M.nblk, kThreads, 0, S->s>>>(
                              P, S->y.p, S->x.p, S->xbar.p, S->c.p, S->lo.p, S->up.p, S->tau, S->st.p, k_lo, check)));
        }
    }
    HIP_CHECK(hipGetLastError());
}

// k is the global iteration index (used only to place the violation reduction).
void launch_dual_ref(nir_session* S, int k_lo, long long k) {
    const DevMatrix& M = S->A;
    if (M.shape.fused_ok) {
        const SegCsr<double> F = M.shape.fused(M.val.p);
        TILE_DISPATCH(M.shape.shape.tile,
                      k_dual_ref<TL><<<F.nblk, kThreads, 0, S->s>>>(F, S->xbar.p, S->y.p, S->rl.p, S->ru.p,
                                                                   S->sigma, S->st.p, S->vpart.p));
    } else {
    for (int p = 0; p < M.shape.nseg; ++p) {
        const Pass<double> P = M.shape.pass(M.val.p, M.part.p, p);
        if (M.shape.nlong[p] > 0) {
            TILE_DISPATCH(M.shape.shape.tile,
                          (k_dual_ref<TL, Pass<double>, 2><<<P.M.nblk, kThreads, 0, S->s>>>(
                              P, S->xbar.p, S->y.p, S->rl.p, S->ru.p, S->sigma, S->st.p, S->vpart.p)));
        } else {
            TILE_DISPATCH(M.shape.shape.tile,
                          (k_dual_ref<TL, Pass<double>, 0><<<P.M.nblk, kThreads, 0, S->s>>>(
                              P, S->xbar.p, S->y.p, S->rl.p, S->ru.p, S->sigma, S->st.p, S->vpart.p)));
        }
    }
    }
    // The stop rule reads the violation of iterations with k % 32 == 0.
    if (S->prm.check_termination && (k % 32) == 0)
        k_vmax_reduce<<<1, kThreads, 0, S->s>>>(S->vpart.p, M.shape.result_blocks(), S->st.p, k_lo);
    HIP_CHECK(hipGetLastError());
}

void enqueue_ref_iteration(nir_session* S, long long k) {
    const int k_lo = static_cast<int>(k - S->base_host);
    launch_primal_ref(S, k_lo);
    launch_dual_ref(S, k_lo, k);
}

// One graph replays 32 iterations starting at a multiple of 32.
void build_ref_graph(nir_session* S) {
    HIP_CHECK(hipStreamBeginCapture(S->s, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < 32; ++i) {
        launch_primal_ref(S, i);
        launch_dual_ref(S, i, i);
    }
    k_advance<<<1, 1, 0, S->s>>>(S->st.p, 32);
    HIP_CHECK(hipStreamEndCapture(S->s, &S->graph));
    HIP_CHECK(hipGraphInstantiate(&S->graph_exec, S->graph, nullptr, nullptr, 0));
}

// Enqueue iterations [k, k + count) of the reference loop.
void enqueue_ref_block(nir_session* S, long long k, long long count) {
    const long long end = k + count;
    while (k < end) {
        if (S->prm.use_graph && (k % 32) == 0 && end - k >= 32) {
            if (!S->graph_exec) build_ref_graph(S);
            if (S->base_host != k) {
                k_set_base<<<1, 1, 0, S->s>>>(S->st.p, static_cast<int>(k));
                HIP_CHECK(hipGetLastError());
                S->base_host = static_cast<int>(k);
            }
            HIP_CHECK(hipGraphLaunch(S->graph_exec, S->s));
            S->base_host += 32;
            k += 32;
        } else {
            enqueue_ref_iteration(S, k);
            ++k;
        }
    }
}

// Iterations [k, k + count) of a small LP in one workgroup.
void enqueue_ref_small(nir_session* S, long long k, long long count) {
    const CsrShape& A = S->A.shape;
    const CsrShape& AT = S->AT.shape;
    auto kern = S->small_mat == 2 ? k_ref_small<2> : S->small_mat == 1 ? k_ref_small<1> : k_ref_small<0>;
    kern<<<1, S->small_threads, S->small_lds, S->s>>>(
        S->m, S->n, AT.rowptr.p, AT.col.p, S->AT.val.p, A.rowptr.p, A.col.p, S->A.val.p, S->x.p, S->xbar.p, S->y.p,
        S->c.p, S->lo.p, S->up.p, S->rl.p, S->ru.p, S->tau, S->sigma, S->st.p, S->vpart.p,
        static_cast<int>(S->vpart.n), static_cast<int>(k), static_cast<int>(count), S->prm.check_termination ? 1 : 0);
    HIP_CHECK(hipGetLastError());
}

// Copy the state word to pinned slot `slot` and mark it pending.
void post_poll(nir_session* S, int slot) {
    HIP_CHECK(hipMemcpyAsync(&S->h_poll[slot], S->st.p, sizeof(RefState), hipMemcpyDeviceToHost, S->s));
    HIP_CHECK(hipEventRecord(S->poll_ev[slot], S->s));
    S->poll_pending[slot] = 1;
}

bool wait_poll(nir_session* S, int slot) {
    if (!S->poll_pending[slot]) return false;
    HIP_CHECK(hipEventSynchronize(S->poll_ev[slot]));
    S->poll_pending[slot] = 0;
    return S->h_poll[slot].done != 0;
}

void finish_ref(nir_session* S) {
    HIP_CHECK(hipStreamSynchronize(S->s));
    for (int i = 0; i < 2; ++i) S->poll_pending[i] = 0;
    RefState h{};
    HIP_CHECK(hipMemcpy(&h, S->st.p, sizeof(RefState), hipMemcpyDeviceToHost));
    if (h.done) {
        S->finished = true;
        S->iters_done = h.iters;
        S->status = 0;
    } else {
        S->iters_done = S->k_next;
    }
}

void run_reference(nir_session* S, long long budget) {
    auto t0 = std::chrono::steady_clock::now();
    while (budget > 0 && !S->finished && S->k_next < S->max_iters) {
        if ((S->k_next % 256) == 0 && S->prm.should_stop && S->prm.should_stop(S->prm.stop_user)) {
            S->status = 4;   // TimeLimit
            S->finished = true;
            break;
        }
        const long long room = 256 - (S->k_next % 256);
        const long long chunk = std::min({budget, room, S->max_iters - S->k_next});
        if (S->small_threads > 0) enqueue_ref_small(S, S->k_next, chunk);
        else enqueue_ref_block(S, S->k_next, chunk);
        S->k_next += chunk;
        budget -= chunk;
        const int slot = static_cast<int>(S->poll_serial++ & 1);
        post_poll(S, slot);
        // Keep one block in flight: wait for the previous one only.
        if (wait_poll(S, slot ^ 1)) break;
    }
    finish_ref(S);
    if (S->k_next >= S->max_iters && !S->finished) S->finished = true;
    S->solve_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Power iteration of the reference (12 rounds of u <- A^T A u / ||A^T A u||),
// with the squared norm summed sequentially on the host as the reference does.
void reference_step_sizes(nir_session* S) {
    std::vector<double> u(static_cast<size_t>(S->n), 1.0);
    DevArray<double> du, dv;
    du.alloc(S->n);
    dv.alloc(S->m);
    du.upload(u.data(), u.size(), S->s);
    double eta = 1.0;
    for (int p = 0; p < 12; ++p) {
        spmv_store(S->A, du.p, dv.p, S->s);
        spmv_store(S->AT, dv.p, du.p, S->s);
        du.download(u.data(), u.size(), S->s);
        HIP_CHECK(hipStreamSynchronize(S->s));
        double n2 = 0.0;
        for (double v : u) n2 += v * v;
        eta = std::sqrt(std::max(1e-12, n2));
        for (double& v : u) v /= eta;
        du.upload(u.data(), u.size(), S->s);
    }
    HIP_CHECK(hipStreamSynchronize(S->s));
    const double smax = std::sqrt(eta);
    S->tau = 0.85 / std::max(1e-6, smax);
    S->sigma = 0.85 / std::max(1e-6, smax);
}

int max_row_length(const std::vector<int>& rowptr) {
    int most = 0;
    for (size_t r = 0; r + 1 < rowptr.size(); ++r) most = std::max(most, rowptr[r + 1] - rowptr[r]);
    return most;
}

// Small LPs run their iterations in one workgroup (k_ref_small) when the
// default layout is in use (no explicit tile or segments: those ask for the
// tiled kernels), x, xbar and y fit its LDS and no row of A or A^T is long
// enough to hold one lane up; with both matrices (and then the costs and
// bounds) in LDS as well when they fit beside the vectors, otherwise only up
// to kSmallGlobalNnz nonzeros (beyond that the tiled kernels' full-chip
// products are faster, measured on MI355X).
constexpr long long kSmallGlobalNnz = 3072;
constexpr int kSmallMaxRow = 512;
constexpr size_t kSmallLdsBytes = 150 * 1024;

void choose_small_path(nir_session* S, const std::vector<int>& a_ptr, const std::vector<int>& at_ptr, bool tuned) {
    S->small_threads = 0;
    S->small_mat = 0;
    if (tuned || S->prm.tile_nnz != 0 || S->prm.seg_width != 0 || S->prm.seg_width_t > 0) return;
    if (S->A.shape.nseg != 1 || S->AT.shape.nseg != 1) return;
    const size_t m = static_cast<size_t>(S->m), n = static_cast<size_t>(S->n), nz = static_cast<size_t>(S->nnz);
    const size_t vec = (2 * n + m) * sizeof(double);
    const size_t ints = m + 1 + n + 1 + 2 * nz;
    const size_t mat = 2 * nz * sizeof(double) + (ints + 1) / 2 * sizeof(double);
    const size_t consts = (3 * n + 2 * m) * sizeof(double);
    int use_mat;
    size_t bytes;
    if (vec + mat + consts <= kSmallLdsBytes) use_mat = 2, bytes = vec + mat + consts;
    else if (vec + mat <= kSmallLdsBytes) use_mat = 1, bytes = vec + mat;
    else if (vec <= kSmallLdsBytes && S->nnz <= kSmallGlobalNnz) use_mat = 0, bytes = vec;
    else return;
    for (int i = 0; i < S->m; ++i)
        if (a_ptr[i + 1] - a_ptr[i] > kSmallMaxRow) return;
    for (int j = 0; j < S->n; ++j)
        if (at_ptr[j + 1] - at_ptr[j] > kSmallMaxRow) return;
    const size_t lds = std::max<size_t>(bytes, 8);
    const void* fn = use_mat == 2   ? reinterpret_cast<const void*>(&k_ref_small<2>)
                     : use_mat == 1 ? reinterpret_cast<const void*>(&k_ref_small<1>)
                                    : reinterpret_cast<const void*>(&k_ref_small<0>);
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    const int most = std::max({S->m, S->n, 64});
    S->small_threads = std::min(1024, (most + 63) / 64 * 64);
    S->small_mat = use_mat;
    S->small_lds = lds;
}

// pdlp mode: LPs up to this many nonzeros (and rows and columns that the
// batch kernel stages in LDS) solve faster as a batch of one than on the
// session's two launches per iteration (measured on MI355X: 1.5-3x up to
// 2,400 nonzeros, 0.9x at 3,700); an explicit layout or unscaling vectors
// keep the session path.
constexpr long long kSmallPdlpNnz = 2500;
constexpr int kSmallPdlpDim = 8192;

bool small_pdlp(const nir_lp* lp, const nir_params* p) {
    if (!lp || !p || p->mode != NIR_MODE_PDLP) return false;
    if (p->tile_nnz != 0 || p->seg_width != 0 || p->seg_width_t > 0 || p->unscale_row || p->unscale_col) return false;
    if (lp->nrow < 1 || lp->ncol < 1 || lp->nrow > kSmallPdlpDim || lp->ncol > kSmallPdlpDim || !lp->col_start)
        return false;
    const long long nnz = lp->col_start[lp->ncol];
    return nnz > 0 && nnz <= kSmallPdlpNnz;
}

// Nonzeros from which tile_nnz = 0 autotunes the layout.
constexpr long long kAutotuneNnz = 1ll << 21;

TileShape shape_from_params(const nir_params& p, bool transpose) {
    TileShape t;
    if (p.tile_nnz == 1024 || p.tile_nnz == 2048 || p.tile_nnz == 4096) t.tile = p.tile_nnz;
    else if (p.tile_nnz != 0 && p.tile_nnz != -1) throw std::invalid_argument("tile_nnz must be -1, 0, 1024, 2048 or 4096");
    if (p.tile_rows > 0) t.rows = std::min(p.tile_rows, kThreads);
    if (p.seg_width < 0) throw std::invalid_argument("seg_width must be >= 0");
    if (p.seg_width_t < -1) throw std::invalid_argument("seg_width_t must be >= -1");
    t.seg_width = (transpose && p.seg_width_t >= 0) ? p.seg_width_t : p.seg_width;
    t.fused = p.seg_fused ? 1 : 0;
    return t;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

NIR_API void nir_params_default(nir_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->mode = NIR_MODE_REFERENCE;
    p->precision = NIR_PREC_FP64;
    p->max_iters = 50000;
    p->tol = 1e-6;
    p->polish_tol = 0.0;
    p->eval_freq = 64;
    p->check_termination = 1;
    p->use_graph = 1;
    p->seg_width_t = -1;
}

NIR_API const char* nir_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

namespace nir {
void record_last_error(const std::string& what) { g_last_error = what; }
}  // namespace nir

extern "C" {

NIR_API int nir_abi_layout(int64_t* out, int32_t cap) {
    const int64_t v[] = {
        static_cast<int64_t>(sizeof(nir_lp)),          static_cast<int64_t>(offsetof(nir_lp, obj_offset)),
        static_cast<int64_t>(sizeof(nir_params)),      static_cast<int64_t>(offsetof(nir_params, seg_fused)),
        static_cast<int64_t>(sizeof(nir_info)),        static_cast<int64_t>(offsetof(nir_info, solve_seconds)),
        static_cast<int64_t>(sizeof(nir_timing)),      static_cast<int64_t>(offsetof(nir_timing, spmtv_passes)),
        static_cast<int64_t>(sizeof(nir_batch_opts)),  static_cast<int64_t>(offsetof(nir_batch_opts, workgroups)),
        static_cast<int64_t>(sizeof(nir_batch_stats)), static_cast<int64_t>(offsetof(nir_batch_stats, solve_seconds)),
        static_cast<int64_t>(sizeof(nir_verify_report)), static_cast<int64_t>(offsetof(nir_verify_report, passed)),
    };
    const int32_t n = static_cast<int32_t>(sizeof(v) / sizeof(v[0]));
    if (!out || cap < n) return n;
    for (int32_t i = 0; i < n; ++i) out[i] = v[i];
    return n;
}

NIR_API int nir_ctx_create(int device, nir_ctx** out) {
    if (!out) return -1;
    *out = nullptr;
    try {
        int count = 0;
        HIP_CHECK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw std::runtime_error("no HIP device " + std::to_string(device));
        HIP_CHECK(hipSetDevice(device));
        auto* c = new nir_ctx;
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw HipError(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        roctxNameHipStream("nirnaya engine stream", c->stream);
        *out = c;
        return 0;
    } catch (const std::exception& ex) {
        g_last_error = ex.what();
        return -1;
    }
}

NIR_API void nir_ctx_destroy(nir_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

NIR_API const char* nir_ctx_error(const nir_ctx* ctx) { return ctx ? ctx->error.c_str() : g_last_error.c_str(); }

NIR_API int nir_session_create(nir_ctx* ctx, const nir_lp* lp, const nir_params* p, nir_session** out) {
    const nir::TraceRange trace("nir_session_create");
    if (!ctx || !out) return -1;
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto S = std::make_unique<nir_session>();
    try {
        check_lp(lp);
        HIP_CHECK(hipSetDevice(ctx->device));
        S->ctx = ctx;
        S->s = ctx->stream;
        if (p) S->prm = *p;
        else nir_params_default(&S->prm);
        if (S->prm.mode != NIR_MODE_REFERENCE && S->prm.mode != NIR_MODE_PDLP)
            throw std::invalid_argument("unknown mode");
        S->m = lp->nrow;
        S->n = lp->ncol;
        S->nnz = lp->ncol > 0 ? lp->col_start[lp->ncol] : 0;
        S->obj_offset = lp->obj_offset;
        S->max_iters = std::max<long long>(0, S->prm.max_iters);
        // The reference loop counts iterations in its index type (int32,
        // src/core/types.hpp:12), and so does the device state here.
        if (S->prm.mode == NIR_MODE_REFERENCE && S->max_iters > INT32_MAX)
            throw std::invalid_argument("reference mode: max_iters above 2^31 - 1");
        const TileShape base = shape_from_params(S->prm, false);
        const TileShape base_t = shape_from_params(S->prm, true);

        // CSR(A^T) is the CSC of A; CSR(A) by a counting transpose that keeps
        // column indices ascending within each row.
        const int m = S->m, n = S->n;
        std::vector<int> at_ptr(static_cast<size_t>(n) + 1);
        for (int j = 0; j <= n; ++j) at_ptr[j] = n > 0 ? static_cast<int>(lp->col_start[j]) : 0;
        if (n == 0) at_ptr.assign(1, 0);
        std::vector<int> a_ptr, a_col;
        std::vector<double> a_val;
        csr_from_csc(lp, a_ptr, a_col, a_val);
        // tile_nnz = -1 autotunes; 0 autotunes only large matrices (where a
        // better layout repays the autotuner's few hundred milliseconds) and
        // keeps the default layout otherwise.
        const bool tune = S->prm.tile_nnz == -1 ||
                          (S->prm.tile_nnz == 0 && S->prm.seg_width == 0 && S->nnz >= kAutotuneNnz);
        // pdlp mode launches one kernel per segment (no fused layout) and
        // iterates in FP32 in mixed precision: time what will run.
        const bool pdlp = S->prm.mode == NIR_MODE_PDLP;
        const bool fp32 = pdlp && S->prm.precision == NIR_PREC_MIXED;
        // pdlp mode sums rows of more than kLongRow nonzeros in the chunked
        // order (spmv.hpp), which a column-segmented layout would cut; a
        // matrix with such rows keeps one segment.
        const bool seg_a = !pdlp || max_row_length(a_ptr) <= kLongRow;
        const bool seg_t = !pdlp || max_row_length(at_ptr) <= kLongRow;
        TileShape ta = tune ? autotune(m, n, a_ptr, a_col.data(), a_val.data(), S->s, fp32, !pdlp, seg_a, pdlp) : base;
        TileShape tt = tune ? autotune(n, m, at_ptr, lp->row_index, lp->value, S->s, fp32, !pdlp, seg_t, pdlp) : base_t;
        if (!seg_a) ta.seg_width = 0, ta.fused = 0;
        if (!seg_t) tt.seg_width = 0, tt.fused = 0;
        // ... and splits each such row over workgroups (pieces of whole
        // groups; the chunked order is kept, so the split changes no bits).
        ta.split = tt.split = pdlp ? 1 : 0;
        upload_matrix(S->A, m, n, a_ptr, a_col.data(), a_val.data(), ta, S->s);
        upload_matrix(S->AT, n, m, at_ptr, lp->row_index, lp->value, tt, S->s);

        S->h_cost.assign(lp->cost, lp->cost + n);
        if (S->prm.mode == NIR_MODE_PDLP) {
            S->pdlp.reset(pdlp_create());
            pdlp_setup(*S->pdlp, S->s, lp, S->prm, &S->A, &S->AT);
        } else {
            S->c.alloc(n); S->c.upload(lp->cost, n, S->s);
            S->lo.alloc(n); S->lo.upload(lp->collo, n, S->s);
            S->up.alloc(n); S->up.upload(lp->colup, n, S->s);
            S->rl.alloc(m); S->rl.upload(lp->rowlo, m, S->s);
            S->ru.alloc(m); S->ru.upload(lp->rowup, m, S->s);
            // Start point (pdhg_solver.hpp:47-57) unless a warm start is given.
            std::vector<double> x0(static_cast<size_t>(n), 0.0), y0(static_cast<size_t>(m), 0.0);
            for (int j = 0; j < n; ++j) {
                const double l = lp->collo[j], u = lp->colup[j];
                if (S->prm.x0) x0[j] = S->prm.x0[j];
                else if (host_finite(l) && host_finite(u)) x0[j] = 0.5 * (l + u);
                else if (host_finite(l)) x0[j] = std::max(0.0, l);
                else if (host_finite(u)) x0[j] = std::min(0.0, u);
            }
            if (S->prm.y0) y0.assign(S->prm.y0, S->prm.y0 + m);
            S->x.alloc(n); S->x.upload(x0.data(), n, S->s);
            S->xbar.alloc(n); S->xbar.upload(x0.data(), n, S->s);
            S->y.alloc(m); S->y.upload(y0.data(), m, S->s);
            S->vpart.alloc(static_cast<size_t>(S->A.shape.result_blocks()));
            {
                std::vector<double> zeros(S->vpart.n, 0.0);
                S->vpart.upload(zeros.data(), zeros.size(), S->s);
            }
            S->st.alloc(1);
            RefState init{};
            S->st.upload(&init, 1, S->s);
            HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&S->h_poll), 2 * sizeof(RefState), hipHostMallocDefault));
            std::memset(S->h_poll, 0, 2 * sizeof(RefState));
            for (auto& e : S->poll_ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            reference_step_sizes(S.get());
            HIP_CHECK(hipStreamSynchronize(S->s));
            choose_small_path(S.get(), a_ptr, at_ptr, tune);
        }
        S->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = S.release();
        return 0;
    } catch (const std::exception& ex) {
        ctx->error = ex.what();
        g_last_error = ex.what();
        // Work already queued on the stream (uploads, autotune products, the
        // power iteration) still reads the buffers S owns: drain it first.
        if (S->s) (void)hipStreamSynchronize(S->s);
        return -1;
    }
}

NIR_API int nir_session_run(nir_session* S, int64_t iters) {
    const nir::TraceRange trace("nir_session_run");
    if (!S) return -1;
    try {
        HIP_CHECK(hipSetDevice(S->ctx->device));
        if (S->pdlp) pdlp_run(*S->pdlp, S->s, S->prm, iters);
        else run_reference(S, iters);
        return 0;
    } catch (const std::exception& ex) {
        S->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_session_kkt(nir_session* S, double* out5) {
    if (!S) return -1;
    try {
        if (!out5) throw std::invalid_argument("missing output");
        if (S->pdlp) throw std::invalid_argument("pdlp sessions report their relative KKT terms in nir_info");
        HIP_CHECK(hipSetDevice(S->ctx->device));
        kkt_device(S->s, S->A.shape, S->A.val.p, S->AT.shape, S->AT.val.p, S->x.p, S->y.p, S->c.p, S->lo.p,
                   S->up.p, S->rl.p, S->ru.p, S->obj_offset, out5);
        return 0;
    } catch (const std::exception& ex) {
        S->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_session_solve(nir_session* S) {
    if (!S) return -1;
    return nir_session_run(S, INT64_MAX / 2);
}

NIR_API int nir_session_result(nir_session* S, double* x, double* y, nir_info* info) {
    if (!S) return -1;
    try {
        HIP_CHECK(hipSetDevice(S->ctx->device));
        if (S->pdlp) return pdlp_result(*S->pdlp, S->s, x, y, info);
        std::vector<double> hx(static_cast<size_t>(S->n));
        S->x.download(hx.data(), hx.size(), S->s);
        if (y) S->y.download(y, static_cast<size_t>(S->m), S->s);
        std::vector<double> vp(S->vpart.n);
        S->vpart.download(vp.data(), vp.size(), S->s);
        HIP_CHECK(hipStreamSynchronize(S->s));
        if (x) std::copy(hx.begin(), hx.end(), x);
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->status = S->status;
            info->iterations = S->iters_done;
            // The reference reports c^T x + offset of the final x in double-double.
            DDHost acc(S->obj_offset);
            for (int j = 0; j < S->n; ++j)
                if (S->h_cost[j] != 0.0 && hx[j] != 0.0) acc.add_prod(S->h_cost[j], hx[j]);
            info->objective = acc.value();
            info->tau = S->tau;
            info->sigma = S->sigma;
            info->primal_weight = 1.0;
            double vmax = 0.0;   // of the last iteration's A xbar
            for (double v : vp) vmax = std::max(vmax, v);
            info->max_primal_viol = S->iters_done > 0 ? vmax : 0.0;
            info->setup_seconds = S->setup_seconds;
            info->solve_seconds = S->solve_seconds;
        }
        return 0;
    } catch (const std::exception& ex) {
        S->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_session_time(nir_session* S, int64_t iters, nir_timing* out) {
    if (!S || !out || iters <= 0) return -1;
    try {
        HIP_CHECK(hipSetDevice(S->ctx->device));
        std::memset(out, 0, sizeof(*out));
        if (S->pdlp) return pdlp_time(*S->pdlp, S->s, iters, out);
        if (S->finished) throw std::runtime_error("session already finished");
        const int m = S->m, n = S->n;
        const double nnz = static_cast<double>(S->nnz);
        out->spmv_bytes = 12.0 * nnz + 4.0 * (m + 1) + 8.0 * n + 32.0 * m;
        out->spmtv_bytes = 12.0 * nnz + 4.0 * (n + 1) + 8.0 * m + 48.0 * n;
        out->iter_bytes = out->spmv_bytes + out->spmtv_bytes;
        // Sweep 1, the production path (32-iteration graph replays), between
        // one pair of events: total_ms.
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipStreamSynchronize(S->s));
        HIP_CHECK(hipEventRecord(e0, S->s));
        enqueue_ref_block(S, S->k_next, iters);
        S->k_next += iters;
        HIP_CHECK(hipEventRecord(e1, S->s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        out->total_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        // Sweep 2, as many iterations launched one by one with an event pair
        // around each product (all its segment passes): spmv_ms and spmtv_ms
        // are the summed product durations of this sweep. The events skip the
        // system-scope fence, which only host visibility of memory needs.
        const long long chunk = 64;
        std::vector<hipEvent_t> ev(static_cast<size_t>(4 * chunk));
        for (auto& e : ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        long long k = S->k_next;
        for (long long t0 = 0; t0 < iters; t0 += chunk) {
            const long long cnt = std::min(chunk, iters - t0);
            for (long long t = 0; t < cnt; ++t, ++k) {
                hipEvent_t* e = ev.data() + 4 * t;
                const int k_lo = static_cast<int>(k - S->base_host);
                HIP_CHECK(hipEventRecord(e[0], S->s));
                launch_primal_ref(S, k_lo);
                HIP_CHECK(hipEventRecord(e[1], S->s));
                HIP_CHECK(hipEventRecord(e[2], S->s));
                launch_dual_ref(S, k_lo, k);
                HIP_CHECK(hipEventRecord(e[3], S->s));
            }
            HIP_CHECK(hipStreamSynchronize(S->s));
            for (long long t = 0; t < cnt; ++t) {
                hipEvent_t* e = ev.data() + 4 * t;
                HIP_CHECK(hipEventElapsedTime(&ms, e[0], e[1]));
                out->spmtv_ms += ms;
                HIP_CHECK(hipEventElapsedTime(&ms, e[2], e[3]));
                out->spmv_ms += ms;
            }
        }
        for (auto& e : ev) HIP_CHECK(hipEventDestroy(e));
        out->iterations = iters;
        out->spmv_passes = S->A.shape.fused_ok ? 1 : S->A.shape.nseg;
        out->spmtv_passes = S->AT.shape.fused_ok ? 1 : S->AT.shape.nseg;
        S->k_next = k;
        finish_ref(S);
        return 0;
    } catch (const std::exception& ex) {
        S->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_session_shape(nir_session* S, int32_t* out8) {
    if (!S || !out8) return -1;
    out8[0] = S->small_threads > 0 ? 0 : S->A.shape.shape.tile;
    out8[1] = S->A.shape.shape.rows;
    out8[2] = S->A.shape.nseg;
    out8[3] = S->A.shape.fused_ok ? 1 : 0;
    out8[4] = S->AT.shape.shape.tile;
    out8[5] = S->AT.shape.shape.rows;
    out8[6] = S->AT.shape.nseg;
    out8[7] = S->AT.shape.fused_ok ? 1 : 0;
    return 0;
}

NIR_API int nir_session_split(nir_session* S, int32_t* out4) {
    if (!S || !out4) return -1;
    out4[0] = S->A.shape.nsplit;
    out4[1] = S->A.shape.npieces;
    out4[2] = S->AT.shape.nsplit;
    out4[3] = S->AT.shape.npieces;
    return 0;
}

NIR_API int nir_calibrate_copy(nir_ctx* ctx, int64_t bytes, double* ms) {
    if (!ctx || bytes <= 0) return -1;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        const long long n = bytes / static_cast<long long>(sizeof(double4));
        DevArray<double4> a, b;
        a.alloc(n);
        b.alloc(n);
        HIP_CHECK(hipMemsetAsync(a.p, 0, n * sizeof(double4), ctx->stream));
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipEventRecord(e0, ctx->stream));
        k_copy_calibrate<<<2048, 256, 0, ctx->stream>>>(a.p, b.p, n);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipEventRecord(e1, ctx->stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float t = 0.f;
        HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        if (ms) *ms = t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return 0;
    } catch (const std::exception& ex) {
        ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_session_gather_floor(nir_session* S, int32_t which, int32_t reps, double* us) {
    if (!S) return -1;
    try {
        if (S->pdlp) throw std::invalid_argument("the access floor is measured on reference-mode sessions");
        if (which != 0 && which != 1) throw std::invalid_argument("which must be 0 (A xbar) or 1 (A^T y)");
        if (reps < 1 || !us) throw std::invalid_argument("reps must be >= 1 and us non-null");
        HIP_CHECK(hipSetDevice(S->ctx->device));
        const DevMatrix& M = which == 0 ? S->A : S->AT;
        const double* v = which == 0 ? S->xbar.p : S->y.p;
        constexpr int kBlocks = 2048;
        DevArray<double> out;
        out.alloc(static_cast<size_t>(kBlocks) * 256);
        auto launch = [&]() {
            // One launch per column segment, as the product makes its passes.
            for (int g = 0; g < M.shape.nseg; ++g)
                k_gather_floor<<<kBlocks, 256, 0, S->s>>>(M.shape.col.p, M.val.p, v, M.shape.seg_nz[g],
                                                          M.shape.seg_nz[g + 1], out.p);
            HIP_CHECK(hipGetLastError());
        };
        launch();   // warm caches and code
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipEventRecord(e0, S->s));
        for (int r = 0; r < reps; ++r) launch();
        HIP_CHECK(hipEventRecord(e1, S->s));
        HIP_CHECK(hipEventSynchronize(e1));
        float t = 0.f;
        HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *us = static_cast<double>(t) * 1e3 / reps;
        return 0;
    } catch (const std::exception& ex) {
        S->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_calibrate_gather(nir_ctx* ctx, int64_t n, int32_t len, int32_t with_values, int32_t blocks,
                                 double* us) {
    if (!ctx) return -1;
    try {
        if (n < 1 || len < 1 || len > n || blocks < 1 || blocks > 65536 || !us || with_values < 0 || with_values > 2)
            throw std::invalid_argument("need 1 <= len <= n, 1 <= blocks <= 65536, mode 0..2 and an output");
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DevArray<int> idx;
        DevArray<double> val, x, out;
        idx.alloc(static_cast<size_t>(n));
        val.alloc(static_cast<size_t>(n));
        x.alloc(static_cast<size_t>(len));
        out.alloc(static_cast<size_t>(blocks) * 256);
        k_fill_gather_probe<<<2048, 256, 0, s>>>(idx.p, val.p, x.p, n, len);
        auto launch = [&]() {
            if (with_values == 2) k_gather_probe<2><<<blocks, 256, 0, s>>>(idx.p, val.p, x.p, n, out.p);
            else if (with_values) k_gather_probe<1><<<blocks, 256, 0, s>>>(idx.p, val.p, x.p, n, out.p);
            else k_gather_probe<0><<<blocks, 256, 0, s>>>(idx.p, val.p, x.p, n, out.p);
        };
        launch();
        launch();
        HIP_CHECK(hipGetLastError());
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        constexpr int kReps = 10;
        HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < kReps; ++r) launch();
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float t = 0.f;
        HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *us = static_cast<double>(t) * 1e3 / kReps;
        return 0;
    } catch (const std::exception& ex) {
        ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_session_scaling(nir_session* S, double* R, double* C) {
    if (!S) return -1;
    try {
        HIP_CHECK(hipSetDevice(S->ctx->device));
        if (!S->pdlp) throw std::invalid_argument("scaling factors exist in pdlp mode only");
        pdlp_scaling(*S->pdlp, S->s, R, C);
        return 0;
    } catch (const std::exception& ex) {
        S->ctx->error = ex.what();
        return -1;
    }
}

NIR_API void nir_session_destroy(nir_session* S) {
    if (!S) return;
    (void)hipSetDevice(S->ctx->device);
    (void)hipStreamSynchronize(S->s);
    delete S;
}

NIR_API int nir_pdhg_solve(nir_ctx* ctx, const nir_lp* lp, const nir_params* p, double* x, double* y,
                           nir_info* info) {
    const nir::TraceRange trace("nir_pdhg_solve");
    if (small_pdlp(lp, p)) {
        // A small LP in pdlp mode runs as a batch of one: a persistent
        // workgroup iterates it from LDS-staged operands instead of two
        // launches per iteration, and returns exactly what the session path
        // returns (the batch engine follows the single-LP trajectory bit for
        // bit). 256 iterations per launch keep should_stop's cadence.
        nir_batch_opts o;
        nir_batch_opts_default(&o);
        o.threads = std::max(lp->nrow, lp->ncol) > 256 ? 1024 : 256;
        o.quantum = 256 / std::max(1, p->eval_freq > 0 ? p->eval_freq : 64);
        if (o.quantum < 1) o.quantum = 1;
        o.rounds_per_graph = 1;
        nir_params q = *p;
        q.x0 = nullptr;
        q.y0 = nullptr;
        const double* x0s[1] = {p->x0};
        const double* y0s[1] = {p->y0};
        nir_batch* B = nullptr;
        if (nir_batch_create_warm(ctx, &lp, 1, x0s, y0s, &q, &o, &B) != 0) return -1;
        int rc = nir_batch_solve(B);
        if (rc == 0) rc = nir_batch_result(B, 0, x, y, info);
        nir_batch_destroy(B);
        return rc;
    }
    nir_session* S = nullptr;
    if (nir_session_create(ctx, lp, p, &S) != 0) return -1;
    int rc = nir_session_solve(S);
    if (rc == 0) rc = nir_session_result(S, x, y, info);
    nir_session_destroy(S);
    return rc;
}

}  // extern "C"
