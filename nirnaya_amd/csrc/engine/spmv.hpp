// CSR-stream sparse products for gfx950.
//
// A matrix is cut into row blocks: each block holds at most `tile_rows`
// (<= 256) consecutive rows whose nonzeros fit one LDS tile of TILE values,
// or a single row longer than TILE. One 256-lane workgroup (four wave64s)
// handles one block:
//   1. stream the block's nonzeros with coalesced (non-temporal) loads of the
//      index and value arrays, gather the dense operand, and park the products
//      in LDS;
//   2. one lane per row sums its products from LDS in ascending column order,
//      starting from 0.0 - the same order and the same rounding as a
//      sequential CSR/CSC dot product, so results are bit-identical to it;
//   3. the caller's epilogue consumes the row sum (fused vector update).
// Rows longer than TILE use strided lane partials and a fixed-order LDS tree.
#pragma once
#include <hip/hip_runtime.h>

namespace nir {
namespace gpu {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

template <typename T>
struct Csr {
    int nrows;
    int ncols;
    const int* rowptr;   // nrows + 1
    const int* col;      // nnz
    const T* val;        // nnz
    int nblk;            // number of row blocks (>= 1)
    const int* blk;      // nblk + 1 row boundaries
};

__device__ __forceinline__ bool finite_bound(double b) { return __builtin_isfinite(b) && fabs(b) < 1e20; }
__device__ __forceinline__ double amax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double amin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ float amaxf(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float aminf(float a, float b) { return (b < a) ? b : a; }

template <typename T>
__device__ __forceinline__ T load_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

// Row sums of A * v for the block `b`; calls on_row(row, sum) once per row.
template <int TILE, typename T, typename V, typename F>
__device__ __forceinline__ void stream_rows(const Csr<T>& A, const V* __restrict__ v, T* prod, int b, F&& on_row) {
    static_assert(TILE % kThreads == 0, "tile must be a multiple of the workgroup");
    const int tid = threadIdx.x;
    const int r0 = A.blk[b], r1 = A.blk[b + 1];
    const int p0 = A.rowptr[r0], p1 = A.rowptr[r1];
    const int cnt = p1 - p0;
    if (cnt > TILE) {
        // One long row: lane partials then a fixed-order tree.
        T s = T(0);
        for (int p = p0 + tid; p < p1; p += kThreads) s += load_stream(A.val + p) * static_cast<T>(v[load_stream(A.col + p)]);
        prod[tid] = s;
        __syncthreads();
        for (int w = kThreads / 2; w > 0; w >>= 1) {
            if (tid < w) prod[tid] += prod[tid + w];
            __syncthreads();
        }
        if (tid == 0) on_row(r0, prod[0]);
        return;
    }
    constexpr int PER = TILE / kThreads;
    int ci[PER];
    T vi[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int q = tid + u * kThreads;
        if (q < cnt) {
            ci[u] = load_stream(A.col + p0 + q);
            vi[u] = load_stream(A.val + p0 + q);
        }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int q = tid + u * kThreads;
        if (q < cnt) prod[q] = vi[u] * static_cast<T>(v[ci[u]]);
    }
    __syncthreads();
    const int r = r0 + tid;
    if (r < r1) {
        const int a = A.rowptr[r] - p0, e = A.rowptr[r + 1] - p0;
        T s = T(0);
        for (int q = a; q < e; ++q) s += prod[q];
        on_row(r, s);
    }
}

// Max of nonnegative values over the workgroup; result valid in thread 0.
__device__ __forceinline__ double block_max(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = amax(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kWaves; ++w) v = amax(v, red[w]);
    return v;
}

}  // namespace gpu
}  // namespace nir
