// Double-double solution certifier on the device: the checks of the
// reference's verify_solution (src/io/sol_io.hpp:25-106) for a continuous LP
// and its objective evaluation (src/model/model.hpp:73-98), with the same
// arithmetic as the host checker:
//   * bound violation max(0, l - x) / (1 + |l|) and max(0, x - u) / (1 + |u|)
//     for finite bounds (|b| < 1e20);
//   * row activity r_i = sum_j a_ij x_j in double-double, accumulated in
//     ascending column order and skipping x_j == 0 (two-product + two-sum, the
//     host DD::add_prod), row violation relative to 1 + |bound|;
//   * objective offset + sum c_j x_j in double-double (skipping zeros), and
//     its relative difference from the claimed objective;
//   * pass iff both violations <= tol and the objective error <= 1e-5.
// One lane per row walks its CSR row sequentially, so each activity is the
// host's value bit for bit; maxima keep the first (index, lower before upper)
// attaining them, as the host's strict running maximum does.
//
// The same file evaluates the relative KKT error of a point (kkt_device) for
// reference-mode sessions, with the definitions of the restarted engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/nirnaya/pdhg.h"
#include "ctx.hpp"
#include "devmat.hpp"
#include "lpcheck.hpp"
#include "pdlp.hpp"
#include "pdlp_core.hpp"
#include "trace.hpp"
#include "verify.hpp"

namespace nir {
namespace gpu {
namespace verify {

struct MaxAt {
    double v;
    int at;
};

__device__ __forceinline__ MaxAt better(MaxAt a, MaxAt b) {
    // larger violation wins; on ties the smaller index (the host's first)
    if (b.v > a.v || (b.v == a.v && b.at >= 0 && (a.at < 0 || b.at < a.at))) return b;
    return a;
}

__device__ __forceinline__ MaxAt block_best(MaxAt m, MaxAt* sh) {
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] = better(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    return sh[0];
}

// Row activities and row violations: one lane per row.
__global__ __launch_bounds__(256) void k_verify_rows(int m, const int* __restrict__ rp, const int* __restrict__ col,
                                                     const double* __restrict__ val, const double* __restrict__ x,
                                                     const double* __restrict__ rl, const double* __restrict__ ru,
                                                     MaxAt* part) {
    __shared__ MaxAt sh[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    MaxAt best{0.0, -1};
    if (i < m) {
        dd act = dd_zero();
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            const double xj = x[col[p]];
            if (xj != 0.0) act = dd_add(act, dd_prod(val[p], xj));
        }
        const double r = dd_val(act);
        // `at` = 2 index + side (0 lower, 1 upper): the host checks the lower
        // bound first and keeps the first strict maximum.
        if (finite_bound(rl[i])) best = better(best, MaxAt{amax(0.0, rl[i] - r) / (1.0 + fabs(rl[i])), 2 * i});
        if (finite_bound(ru[i])) best = better(best, MaxAt{amax(0.0, r - ru[i]) / (1.0 + fabs(ru[i])), 2 * i + 1});
    }
    const MaxAt b = block_best(best, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// Bound violations, non-finite entries and the objective's dd partials.
__global__ __launch_bounds__(256) void k_verify_cols(int n, const double* __restrict__ x,
                                                     const double* __restrict__ lo, const double* __restrict__ up,
                                                     const double* __restrict__ c, MaxAt* part, dd* obj_part,
                                                     int* nonfinite) {
    __shared__ MaxAt sh[256];
    __shared__ dd so[256];
    const int j = blockIdx.x * 256 + threadIdx.x;
    MaxAt best{0.0, -1};
    dd o = dd_zero();
    if (j < n) {
        const double v = x[j];
        if (!isfinite(v)) {
            atomicMin(nonfinite, j);
        } else {
            if (finite_bound(lo[j])) best = better(best, MaxAt{amax(0.0, lo[j] - v) / (1.0 + fabs(lo[j])), 2 * j});
            if (finite_bound(up[j])) best = better(best, MaxAt{amax(0.0, v - up[j]) / (1.0 + fabs(up[j])), 2 * j + 1});
            if (c[j] != 0.0 && v != 0.0) o = dd_prod(c[j], v);
        }
    }
    const MaxAt b = block_best(best, sh);
    so[threadIdx.x] = o;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) so[threadIdx.x] = dd_add(so[threadIdx.x], so[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = b;
        obj_part[blockIdx.x] = so[0];
    }
}

// ------------------------------------------------------------------ relative KKT

// Fixed-order tree over the block of four dd terms per lane; lane 0 writes
// the block's four sums to part[4 b ..].
__device__ __forceinline__ void block_sum4(dd (&t)[4], dd (*sh)[256], dd* part) {
    for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = t[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = dd_add(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = sh[k][0];
}

// Row terms: 0 primal residual^2, 1 ||b||^2 term, 2 dual sign violation^2,
// 3 dual-objective row term.
__global__ __launch_bounds__(256) void k_kkt_rows(int m, int nseg, const int* __restrict__ rp,
                                                  const int* __restrict__ col, const double* __restrict__ val,
                                                  const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ rl, const double* __restrict__ ru,
                                                  dd* part) {
    __shared__ dd sh[4][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    dd t[4] = {dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    if (i < m) {
        const double ax = dd_val(row_dot_dd(m, nseg, rp, col, val, x, i));
        const double l = rl[i], u = ru[i];
        const bool lf = finite_bound(l), uf = finite_bound(u);
        double proj = ax;
        if (lf) proj = amax(l, proj);
        if (uf) proj = amin(u, proj);
        const double r = ax - proj;
        t[0] = dd_prod(r, r);
        double b = 0.0;
        if (lf) b = fabs(l);
        if (uf) b = amax(b, fabs(u));
        t[1] = dd_prod(b, b);
        double yp = y[i];
        if (!lf) yp = amin(yp, 0.0);
        if (!uf) yp = amax(yp, 0.0);
        const double yv = y[i] - yp;
        t[2] = dd_prod(yv, yv);
        if (lf && yp > 0.0) t[3] = dd_prod(l, yp);
        if (uf && yp < 0.0) t[3] = dd_add(t[3], dd_prod(u, yp));
    }
    block_sum4(t, sh, part);
}

// Column terms: 0 dual residual^2, 1 ||c||^2 term, 2 primal objective term,
// 3 dual-objective column term.
__global__ __launch_bounds__(256) void k_kkt_cols(int n, int nseg, const int* __restrict__ rp,
                                                  const int* __restrict__ col, const double* __restrict__ val,
                                                  const double* __restrict__ y, const double* __restrict__ x,
                                                  const double* __restrict__ c, const double* __restrict__ lo,
                                                  const double* __restrict__ up, dd* part) {
    __shared__ dd sh[4][256];
    const int j = blockIdx.x * 256 + threadIdx.x;
    dd t[4] = {dd_zero(), dd_zero(), dd_zero(), dd_zero()};
    if (j < n) {
        const double aty = dd_val(row_dot_dd(n, nseg, rp, col, val, y, j));
        const double l = lo[j], u = up[j], cj = c[j], xj = x[j];
        const bool lf = finite_bound(l), uf = finite_bound(u);
        const double lam = cj - aty;
        double lp = lam;
        if (!lf) lp = amin(lp, 0.0);
        if (!uf) lp = amax(lp, 0.0);
        const double r = lam - lp;
        t[0] = dd_prod(r, r);
        t[1] = dd_prod(cj, cj);
        if (xj != 0.0) t[2] = dd_prod(cj, xj);
        if (lf && lp > 0.0) t[3] = dd_prod(l, lp);
        if (uf && lp < 0.0) t[3] = dd_add(t[3], dd_prod(u, lp));
    }
    block_sum4(t, sh, part);
}

// ------------------------------------------------------------------ allocation-free certifier

// Largest nrow / ncol of nir_verify_solution_noalloc: at most kNoAllocBlocks
// 256-lane blocks per pass, so its block partials fit the static scratch.
constexpr int kNoAllocBlocks = 65536;
constexpr long long kNoAllocMaxDim = static_cast<long long>(kNoAllocBlocks) * 256;

// Result of the final reduction, read back by the host.
struct VerifyOut {
    double max_row, max_bound, obj;
    int row_at, col_at, nonfinite, pad;
};

// Scratch of the allocation-free entry point: row maxima, column maxima,
// objective dd partials (one slot per block each), the non-finite column
// index and the reduced result. One copy per GPU, shared by every call on it.
constexpr size_t kScratchRowOff = 0;
constexpr size_t kScratchColOff = kScratchRowOff + sizeof(MaxAt) * kNoAllocBlocks;
constexpr size_t kScratchObjOff = kScratchColOff + sizeof(MaxAt) * kNoAllocBlocks;
constexpr size_t kScratchNfOff = kScratchObjOff + sizeof(dd) * kNoAllocBlocks;
constexpr size_t kScratchOutOff = kScratchNfOff + 256;
constexpr size_t kScratchBytes = kScratchOutOff + 256;
static_assert(sizeof(VerifyOut) <= 256, "result slot");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void k_set_int(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Reduces the block partials in the order the host loop of
// nir_verify_solution does: the largest positive violation, the first block
// attaining it; the objective as offset + partials added in block order with
// the same double-double sum.
__global__ __launch_bounds__(256) void k_verify_final(int gr, int gc, const MaxAt* __restrict__ prow,
                                                      const MaxAt* __restrict__ pcol, const dd* __restrict__ pobj,
                                                      const int* __restrict__ nonfinite, double offset,
                                                      VerifyOut* out) {
    __shared__ MaxAt sh[256];
    __shared__ int shb[256];
    auto reduce = [&](const MaxAt* part, int count, double& v, int& at) {
        MaxAt best{0.0, -1};
        int bb = -1;
        for (int b = threadIdx.x; b < count; b += 256)
            if (part[b].at >= 0 && part[b].v > best.v) {
                best = part[b];
                bb = b;
            }
        sh[threadIdx.x] = best;
        shb[threadIdx.x] = bb;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) {
                const MaxAt o = sh[threadIdx.x + w];
                const int ob = shb[threadIdx.x + w];
                const MaxAt c = sh[threadIdx.x];
                const int cb = shb[threadIdx.x];
                if (o.v > c.v || (o.v == c.v && ob >= 0 && (cb < 0 || ob < cb))) {
                    sh[threadIdx.x] = o;
                    shb[threadIdx.x] = ob;
                }
            }
            __syncthreads();
        }
        v = sh[0].v;
        at = shb[0] >= 0 ? sh[0].at : -1;
        __syncthreads();
    };
    double vr, vc;
    int ar, ac;
    reduce(prow, gr, vr, ar);
    reduce(pcol, gc, vc, ac);
    if (threadIdx.x == 0) {
        dd acc{offset, 0.0};
        for (int b = 0; b < gc; ++b) acc = dd_add(acc, pobj[b]);
        out->max_row = vr;
        out->row_at = ar;
        out->max_bound = vc;
        out->col_at = ac;
        out->obj = acc.hi + acc.lo;
        out->nonfinite = *nonfinite;
    }
}

}  // namespace verify

void kkt_device(hipStream_t s, const CsrShape& A, const double* aval, const CsrShape& AT, const double* atval,
                const double* x, const double* y, const double* c, const double* lo, const double* up,
                const double* rl, const double* ru, double offset, double* out) {
    const int m = A.nrows, n = AT.nrows;
    const int gr = std::max(1, (m + 255) / 256), gc = std::max(1, (n + 255) / 256);
    DevArray<dd> prow, pcol;
    prow.alloc(static_cast<size_t>(4) * gr);
    pcol.alloc(static_cast<size_t>(4) * gc);
    verify::k_kkt_rows<<<gr, 256, 0, s>>>(m, A.nseg, A.rowptr.p, A.col.p, aval, x, y, rl, ru, prow.p);
    verify::k_kkt_cols<<<gc, 256, 0, s>>>(n, AT.nseg, AT.rowptr.p, AT.col.p, atval, y, x, c, lo, up, pcol.p);
    HIP_CHECK(hipGetLastError());
    std::vector<dd> hr(static_cast<size_t>(4) * gr), hc(static_cast<size_t>(4) * gc);
    prow.download(hr.data(), hr.size(), s);
    pcol.download(hc.data(), hc.size(), s);
    HIP_CHECK(hipStreamSynchronize(s));
    DDHost rp, bn, rd, cn, P(offset), D(offset);
    for (int b = 0; b < gr; ++b) {
        rp.add(hr[4 * b].hi, hr[4 * b].lo);
        bn.add(hr[4 * b + 1].hi, hr[4 * b + 1].lo);
        rd.add(hr[4 * b + 2].hi, hr[4 * b + 2].lo);
        D.add(hr[4 * b + 3].hi, hr[4 * b + 3].lo);
    }
    for (int b = 0; b < gc; ++b) {
        rd.add(hc[4 * b].hi, hc[4 * b].lo);
        cn.add(hc[4 * b + 1].hi, hc[4 * b + 1].lo);
        P.add(hc[4 * b + 2].hi, hc[4 * b + 2].lo);
        D.add(hc[4 * b + 3].hi, hc[4 * b + 3].lo);
    }
    const double pv = P.value(), dv = D.value();
    out[0] = std::sqrt(rp.value()) / (1.0 + std::sqrt(bn.value()));
    out[1] = std::sqrt(rd.value()) / (1.0 + std::sqrt(cn.value()));
    out[2] = std::fabs(pv - dv) / (1.0 + std::fabs(pv) + std::fabs(dv));
    out[3] = pv;
    out[4] = dv;
}

}  // namespace gpu
}  // namespace nir

using namespace nir::gpu;
using namespace nir::gpu::verify;

namespace {

// Device address of `scratch` on each GPU, looked up once per device with
// that device current. Host code uses only these addresses: the array's name
// in host code is a host-side stand-in, not the device address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

unsigned char* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned null");
        g_scratch[device] = static_cast<unsigned char*>(ptr);
    }
    return g_scratch[device];
}

}  // namespace

extern "C" {

NIR_API int nir_verify_solution(nir_ctx* ctx, const nir_lp* lp, const double* x, double claimed_obj, double tol,
                                nir_verify_report* out) {
    const nir::TraceRange trace("nir_verify_solution");
    if (!ctx || !out) return -1;
    try {
        check_lp(lp);
        if (lp->ncol > 0 && !x) throw std::invalid_argument("missing solution vector");
        if (lp->nrow >= (1 << 30) || lp->ncol >= (1 << 30)) throw std::invalid_argument("LP too large to certify");
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        const int m = lp->nrow, n = lp->ncol;
        std::vector<int> ptr, col;
        std::vector<double> val;
        csr_from_csc(lp, ptr, col, val);
        DevArray<int> dptr, dcol;
        DevArray<double> dval, dx, drl, dru, dlo, dup, dc;
        dptr.alloc(ptr.size());
        dptr.upload(ptr.data(), ptr.size(), s);
        dcol.alloc(col.size());
        dcol.upload(col.data(), col.size(), s);
        dval.alloc(val.size());
        dval.upload(val.data(), val.size(), s);
        dx.alloc(n);
        dx.upload(x, n, s);
        dlo.alloc(n);
        dlo.upload(lp->collo, n, s);
        dup.alloc(n);
        dup.upload(lp->colup, n, s);
        dc.alloc(n);
        dc.upload(lp->cost, n, s);
        drl.alloc(m);
        drl.upload(lp->rowlo, m, s);
        dru.alloc(m);
        dru.upload(lp->rowup, m, s);
        const int gr = std::max(1, (m + 255) / 256), gc = std::max(1, (n + 255) / 256);
        DevArray<MaxAt> prow, pcol;
        DevArray<dd> pobj;
        DevArray<int> nonfinite;
        prow.alloc(gr);
        pcol.alloc(gc);
        pobj.alloc(gc);
        nonfinite.alloc(1);
        const int big = 0x7fffffff;
        nonfinite.upload(&big, 1, s);
        k_verify_rows<<<gr, 256, 0, s>>>(m, dptr.p, dcol.p, dval.p, dx.p, drl.p, dru.p, prow.p);
        k_verify_cols<<<gc, 256, 0, s>>>(n, dx.p, dlo.p, dup.p, dc.p, pcol.p, pobj.p, nonfinite.p);
        HIP_CHECK(hipGetLastError());
        std::vector<MaxAt> hr(gr), hc(gc);
        std::vector<dd> ho(gc);
        int nf = big;
        prow.download(hr.data(), gr, s);
        pcol.download(hc.data(), gc, s);
        pobj.download(ho.data(), gc, s);
        nonfinite.download(&nf, 1, s);
        HIP_CHECK(hipStreamSynchronize(s));
        std::memset(out, 0, sizeof(*out));
        out->worst_row = -1;
        out->worst_col = -1;
        out->nonfinite_col = nf == big ? -1 : nf;
        // Blocks hold consecutive indices, so taking the first strict maximum
        // over blocks in order keeps the smallest index among equal maxima.
        for (int b = 0; b < gr && m > 0; ++b)
            if (hr[b].at >= 0 && hr[b].v > out->max_row_viol) {
                out->max_row_viol = hr[b].v;
                out->worst_row = hr[b].at / 2;
                out->worst_row_upper = hr[b].at & 1;
            }
        for (int b = 0; b < gc && n > 0; ++b)
            if (hc[b].at >= 0 && hc[b].v > out->max_bound_viol) {
                out->max_bound_viol = hc[b].v;
                out->worst_col = hc[b].at / 2;
                out->worst_col_upper = hc[b].at & 1;
            }
        DDHost acc(lp->obj_offset);
        for (int b = 0; b < gc && n > 0; ++b) acc.add(ho[b].hi, ho[b].lo);
        out->computed_obj = acc.value();
        out->obj_rel_err = std::isfinite(claimed_obj)
                               ? std::fabs(out->computed_obj - claimed_obj) / (1.0 + std::fabs(out->computed_obj))
                               : 0.0;
        out->passed = out->nonfinite_col < 0 && out->max_bound_viol <= tol && out->max_row_viol <= tol &&
                              out->obj_rel_err <= 1e-5
                          ? 1
                          : 0;
        return 0;
    } catch (const std::exception& ex) {
        ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_verify_solution_noalloc(nir_ctx* ctx, const nir_dev_lp* lp, const double* x, double claimed_obj,
                                        double tol, nir_verify_report* out) {
    if (!ctx || !out) return -1;
    try {
        if (!lp) throw std::invalid_argument("null LP");
        if (lp->nrow < 0 || lp->ncol < 0) throw std::invalid_argument("negative dimensions");
        if (lp->nrow > kNoAllocMaxDim || lp->ncol > kNoAllocMaxDim)
            throw std::invalid_argument("nir_verify_solution_noalloc: nrow and ncol must be at most 16,777,216 "
                                        "(the static scratch holds 65,536 block partials per pass)");
        if (lp->nrow > 0 && (!lp->row_start || !lp->rowlo || !lp->rowup))
            throw std::invalid_argument("missing device row arrays");
        if (lp->ncol > 0 && (!x || !lp->cost || !lp->collo || !lp->colup))
            throw std::invalid_argument("missing device column arrays or solution vector");
        HIP_CHECK(hipSetDevice(ctx->device));
        unsigned char* base = noalloc_scratch(ctx->device);
        hipStream_t s = ctx->stream;
        const int m = lp->nrow, n = lp->ncol;
        const int gr = std::max(1, (m + 255) / 256), gc = std::max(1, (n + 255) / 256);
        MaxAt* prow = reinterpret_cast<MaxAt*>(base + kScratchRowOff);
        MaxAt* pcol = reinterpret_cast<MaxAt*>(base + kScratchColOff);
        dd* pobj = reinterpret_cast<dd*>(base + kScratchObjOff);
        int* nf = reinterpret_cast<int*>(base + kScratchNfOff);
        VerifyOut* res = reinterpret_cast<VerifyOut*>(base + kScratchOutOff);
        const int big = 0x7fffffff;
        // The partial slots are written in full by the two passes below; only
        // the non-finite index must start fresh (a kernel, not a memset node).
        k_set_int<<<1, 64, 0, s>>>(nf, big);
        k_verify_rows<<<gr, 256, 0, s>>>(m, lp->row_start, lp->col_index, lp->value, x, lp->rowlo, lp->rowup, prow);
        k_verify_cols<<<gc, 256, 0, s>>>(n, x, lp->collo, lp->colup, lp->cost, pcol, pobj, nf);
        k_verify_final<<<1, 256, 0, s>>>(m > 0 ? gr : 0, n > 0 ? gc : 0, prow, pcol, pobj, nf, lp->obj_offset, res);
        HIP_CHECK(hipGetLastError());
        VerifyOut h{};
        HIP_CHECK(hipMemcpyAsync(&h, res, sizeof(h), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        std::memset(out, 0, sizeof(*out));
        out->max_row_viol = h.max_row;
        out->worst_row = h.row_at >= 0 ? h.row_at / 2 : -1;
        out->worst_row_upper = h.row_at >= 0 ? (h.row_at & 1) : 0;
        out->max_bound_viol = h.max_bound;
        out->worst_col = h.col_at >= 0 ? h.col_at / 2 : -1;
        out->worst_col_upper = h.col_at >= 0 ? (h.col_at & 1) : 0;
        out->nonfinite_col = h.nonfinite == big ? -1 : h.nonfinite;
        out->computed_obj = h.obj;
        out->obj_rel_err = std::isfinite(claimed_obj)
                               ? std::fabs(out->computed_obj - claimed_obj) / (1.0 + std::fabs(out->computed_obj))
                               : 0.0;
        out->passed = out->nonfinite_col < 0 && out->max_bound_viol <= tol && out->max_row_viol <= tol &&
                              out->obj_rel_err <= 1e-5
                          ? 1
                          : 0;
        return 0;
    } catch (const std::exception& ex) {
        ctx->error = ex.what();
        return -1;
    }
}

}  // extern "C"
