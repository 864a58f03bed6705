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
//      starting from the row's running sum (0.0 on the first pass) - the same
//      order and the same rounding as a sequential CSR/CSC dot product, so
//      results are bit-identical to it;
//   3. the caller's epilogue consumes the row sum (fused vector update).
// A row longer than TILE is a block of its own and is summed as a chain of
// tiles with the same sequential order.
//
// Column segments: a matrix may be split by column ranges into S segments
// (each a CSR over all rows, stored segment-major in one value/index array).
// A product then runs S passes; pass s gathers only from the s-th slice of
// the dense operand, which stays resident in each XCD's L2, and carries the
// running row sums in a partial buffer. Because segments are taken in
// ascending column order and each pass continues the same sequential sum, the
// result is still bit-identical to the single-pass product.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>

namespace nir {
namespace gpu {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

template <typename T>
struct Csr {
    int nrows;
    int ncols;
    const int* rowptr;   // nrows + 1 (absolute offsets into col / val)
    const int* col;      // nnz
    const T* val;        // nnz
    int nblk;            // number of row blocks (>= 1)
    const int4* desc;    // per block {first row, end row, first nonzero, end nonzero}; a piece
                         // of a split row has end row -(split row index + 1)
    // Rows split over workgroups (pdlp mode only, see stream_rows):
    const int4* lrow;    // per split row {first group slot, groups, pieces, row}
    int* lcount;         // per split row: pieces finished in the current launch
    T* gsum;             // group sums of the split rows
};

// One pass of a (possibly segmented) product: which segment, and where the
// running row sums live between passes.
template <typename T>
struct Pass {
    Csr<T> M;
    T* part;     // running row sums (nrows), unused when first && last
    int first;   // start sums at 0.0
    int last;    // hand the completed sum to the epilogue
};

__device__ __forceinline__ bool finite_bound(double b) { return __builtin_isfinite(b) && fabs(b) < 1e20; }
__device__ __forceinline__ double amax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double amin(double a, double b) { return (b < a) ? b : a; }

template <typename T>
__device__ __forceinline__ T load_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

// LDS product tile layout: one padding slot after every 32 products, so lanes
// summing rows that start ~24 products apart hit different banks.
__host__ __device__ constexpr int lds_slots(int tile) { return tile + tile / 32; }
__device__ __forceinline__ int lds_at(int q) { return q + (q >> 5); }

// No epilogue operands to prefetch.
struct NoPre {};
struct PreNone {
    __device__ __forceinline__ NoPre operator()(int) const { return NoPre{}; }
};

// Row sums of A * v for block `b`. On the last pass on_row(row, sum, pre) is
// called once per row, where pre = prefetch(row) was issued before the tile's
// barrier so the epilogue operands arrive while the products are summed; on
// earlier passes the sums are stored to P.part.
//
// The tile loop is branch-free: lanes past the block's nonzeros load slot 0
// (every array is allocated with at least one element), then use column 0
// and value 0 instead, gather v[0] (the dense operand also has at least one
// element), and store into LDS slots the row phase never reads; so all loads
// and gathers of a lane are in flight together.
// Rows longer than this many nonzeros are summed in the chunked order when
// ORDER == 1 (see stream_rows).
constexpr int kLongRow = 4096;
constexpr int kChunk = 64;        // products per chunk
constexpr int kGroupChunks = 32;  // chunks per group
constexpr int kGroupLen = kChunk * kGroupChunks;   // products per group (2048)

// LONG selects the blocks a launch serves: 2 all, 0 only blocks of short rows
// (the long-row code is then not compiled in, so it adds nothing to the
// kernel's register count; blocks of one row longer than a tile and pieces of
// split rows return at once).
// SEQ: products read ahead of the adds in the one-lane sequential sum of a
// long row (16 in the iteration kernels compiled for layouts with long rows,
// 8 elsewhere, where the registers would cost every layout).
template <int TILE, int ORDER = 0, int LONG = 2, int SEQ = 8, typename T, typename V, typename Pre, typename F>
__device__ __forceinline__ void stream_rows(const Pass<T>& P, const V* __restrict__ v, T* prod, int b,
                                            Pre&& prefetch, F&& on_row) {
    static_assert(TILE % kThreads == 0, "tile must be a multiple of the workgroup");
    static_assert(TILE % kChunk == 0 && TILE / kChunk <= 64, "tiles hold whole chunks");
    const Csr<T>& A = P.M;
    const int tid = threadIdx.x;
    const int4 d = A.desc[b];   // one load instead of blk -> rowptr chained loads
    const int r0 = d.x, r1 = d.y, p0 = d.z, p1 = d.w;
    const int cnt = p1 - p0;
    // ORDER == 1: a row of more than kLongRow nonzeros may be split into
    // pieces of whole groups (below), one workgroup each.
    const bool piece = ORDER == 1 && r1 < 0;
    if (!piece && r1 <= r0) return;   // empty matrix: the one block has no rows
    const bool long_block = piece || cnt > TILE;
    if constexpr (LONG == 0) {
        if (long_block) return;   // the layout has none when LONG == 0 is launched
    } else if (long_block) {
        // One row longer than a tile, or one piece of a split row: a chain of
        // tiles. All lanes park one tile of the row's products in LDS, then
        // lane 0 adds them in column order to the running sum, so the row
        // keeps the sequential order and rounding of a CSR dot product at any
        // length (ORDER == 0, and rows up to kLongRow nonzeros); the chunked
        // order below otherwise. The next tile's index and value loads are
        // issued before the current one is summed.
        constexpr int PER = TILE / kThreads;
        T s = P.first ? T(0) : P.part[r0];   // lane 0's copy is the one used
        int ci[PER];
        T vi[PER];
        auto load = [&](int c0) {
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int q = c0 + tid + u * kThreads;
                const bool in = q < p1;
                const int slot = in ? q : 0;
                const int cl = load_stream(A.col + slot);
                const T vl = load_stream(A.val + slot);
                ci[u] = in ? cl : 0;
                vi[u] = in ? vl : T(0);
            }
        };
        // ORDER == 1 and more than kLongRow nonzeros (pdlp mode; such a row
        // is never cut by column segments, so it starts at p0, or a piece
        // starts on one of its group boundaries): the chunked order. Products are summed from 0.0 in chunks of kChunk
        // consecutive positions, chunk sums from 0.0 in groups of
        // kGroupChunks, group sums from 0.0 into the row sum - each level
        // sequential, so 32 (or 64) lanes sum a tile's chunks at once and
        // lane 0 adds only one value per chunk. Rows up to kLongRow nonzeros
        // keep the sequential order in every mode.
        //
        // A piece of a split row covers whole groups from a group boundary of
        // its row (the row's last piece may end in a partial group). Lane 0
        // stores each completed group sum in the row's slots of gsum instead
        // of adding it; the piece that finishes last (a counter per split
        // row, back to 0 once all pieces counted) adds the row's group sums
        // from 0.0 in group order - the same sum whichever piece is last -
        // and runs the epilogue.
        const bool chunked = ORDER == 1 && (piece || cnt > kLongRow);
        __shared__ T chunk_sum[TILE / kChunk];
        T group = T(0);     // lane 0: the current group's partial sum
        int in_group = 0;   // chunks added to it
        int4 L = make_int4(0, 0, 0, 0);
        int slot_g = 0;     // piece: gsum slot of the current group
        if (piece) {
            L = A.lrow[-r1 - 1];
            slot_g = L.x + (p0 - A.rowptr[L.w]) / kGroupLen;
            s = T(0);
        }
        load(p0);
        for (int c0 = p0; c0 < p1; c0 += TILE) {
            T g[PER];
#pragma unroll
            for (int u = 0; u < PER; ++u) g[u] = static_cast<T>(v[ci[u]]);
#pragma unroll
            for (int u = 0; u < PER; ++u) prod[lds_at(tid + u * kThreads)] = vi[u] * g[u];
            __syncthreads();
            if (c0 + TILE < p1) load(c0 + TILE);
            if (chunked) {
                const int e = p1 - c0 < TILE ? p1 - c0 : TILE;
                const int nch = (e + kChunk - 1) / kChunk;
                if (tid < nch) {
                    const int q1 = (tid + 1) * kChunk < e ? (tid + 1) * kChunk : e;
                    T cs = T(0);
                    for (int q = tid * kChunk; q < q1; ++q) cs += prod[lds_at(q)];
                    chunk_sum[tid] = cs;
                }
                __syncthreads();
                if (tid == 0) {
                    for (int k = 0; k < nch; ++k) {
                        group += chunk_sum[k];
                        if (++in_group == kGroupChunks) {
                            if (piece) A.gsum[slot_g++] = group;
                            else s += group;
                            group = T(0);
                            in_group = 0;
                        }
                    }
                }
            } else if (tid == 0) {
                // The same sequential sum, its LDS reads double-buffered: the
                // next eight products are read while the current eight are
                // added, so the chain of adds, not the LDS latency, sets the
                // pace. (Batches of 16 would hide more of the latency but
                // raise the whole kernel's VGPR count - 74 to 118 for the
                // 2048 tile - and cut the occupancy of the short-row blocks.)
                constexpr int B = SEQ;
                const int e = p1 - c0 < TILE ? p1 - c0 : TILE;
                int q = 0;
                if (e >= B) {
                    T a[B], b2[B];
                    auto read = [&](T(&r)[B], int at) {
#pragma unroll
                        for (int k = 0; k < B; ++k) r[k] = prod[lds_at(at + k)];
                        __builtin_amdgcn_sched_barrier(0);   // issue the reads before the adds
                    };
                    auto add = [&](const T(&r)[B]) {
#pragma unroll
                        for (int k = 0; k < B; ++k) s += r[k];
                    };
                    read(a, 0);
                    for (; q + 3 * B <= e; q += 2 * B) {   // a holds [q, q + B)
                        read(b2, q + B);
                        add(a);
                        read(a, q + 2 * B);
                        add(b2);
                    }
                    if (q + 2 * B <= e) {
                        read(b2, q + B);
                        add(a);
                        add(b2);
                        q += 2 * B;
                    } else {
                        add(a);
                        q += B;
                    }
                }
                for (; q < e; ++q) s += prod[lds_at(q)];
            }
            __syncthreads();
        }
        if (piece) {
            __shared__ int last_piece;
            int* count = A.lcount + (-r1 - 1);
            if (tid == 0) {
                if (in_group > 0) A.gsum[slot_g] = group;
                // Release this piece's group sums, then count it.
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                const int before = __hip_atomic_fetch_add(count, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                last_piece = before == L.z - 1;
                if (last_piece) __hip_atomic_store(count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            if (!last_piece) return;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            // All lanes fetch the group sums into LDS, lane 0 adds them.
            for (int k0 = 0; k0 < L.y; k0 += kThreads) {
                const int e = L.y - k0 < kThreads ? L.y - k0 : kThreads;
                if (tid < e) prod[lds_at(tid)] = A.gsum[L.x + k0 + tid];
                __syncthreads();
                if (tid == 0)
                    for (int q = 0; q < e; ++q) s += prod[lds_at(q)];
                __syncthreads();
            }
        } else if (tid == 0 && chunked && in_group > 0) {
            s += group;
        }
        if (tid == 0) {
            if (P.last) on_row(r0, s, prefetch(r0));
            else P.part[r0] = s;
        }
        return;
    }
    // Row bookkeeping and epilogue operands first: their latency overlaps the tile.
    const int r = r0 + tid;
    const bool has_row = r < r1;
    const int rr = has_row ? r : r0;
    const int ra = A.rowptr[rr] - p0, re = A.rowptr[rr + 1] - p0;
    const T init = (P.first || !has_row) ? T(0) : P.part[rr];
    const auto pre = prefetch(rr);

    constexpr int PER = TILE / kThreads;
    int ci[PER];
    T vi[PER];
    T g[PER];
    if constexpr (std::is_same<T, float>::value) {
        // FP32 products: all gathers issue together once the indices are in
        // (a scheduling barrier keeps them after the index loads; the values
        // are masked only at the products). Measured faster for FP32 and
        // slower for FP64, which keeps the compiler's interleaving below.
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + u * kThreads;
            const bool in = q < cnt;
            const int slot = in ? p0 + q : 0;
            const int cl = load_stream(A.col + slot);
            vi[u] = load_stream(A.val + slot);
            ci[u] = in ? cl : 0;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < PER; ++u) g[u] = static_cast<T>(v[ci[u]]);
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const T w = tid + u * kThreads < cnt ? vi[u] : T(0);
            prod[lds_at(tid + u * kThreads)] = w * g[u];
        }
    } else {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + u * kThreads;
            const bool in = q < cnt;
            const int slot = in ? p0 + q : 0;
            const int cl = load_stream(A.col + slot);
            const T vl = load_stream(A.val + slot);
            ci[u] = in ? cl : 0;
            vi[u] = in ? vl : T(0);
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) g[u] = static_cast<T>(v[ci[u]]);
#pragma unroll
        for (int u = 0; u < PER; ++u) prod[lds_at(tid + u * kThreads)] = vi[u] * g[u];
    }
    __syncthreads();
    if (has_row) {
        // Sequential sum in column order; four LDS reads in flight per step.
        T s = init;
        int q = ra;
        for (; q + 4 <= re; q += 4) {
            const T a0 = prod[lds_at(q)], a1 = prod[lds_at(q + 1)], a2 = prod[lds_at(q + 2)], a3 = prod[lds_at(q + 3)];
            s += a0;
            s += a1;
            s += a2;
            s += a3;
        }
        for (; q < re; ++q) s += prod[lds_at(q)];
        if (P.last) on_row(r, s, pre);
        else P.part[r] = s;
    }
}

// Same without epilogue prefetch.
template <int TILE, int ORDER = 0, int LONG = 2, typename T, typename V, typename F>
__device__ __forceinline__ void stream_rows(const Pass<T>& P, const V* __restrict__ v, T* prod, int b, F&& on_row) {
    stream_rows<TILE, ORDER, LONG>(P, v, prod, b, PreNone{}, [&](int r, T s, NoPre) { on_row(r, s); });
}

// ------------------------------------------------------------------ fused segments

// All column segments of a product in one launch: block b owns rows
// [rows[b].x, rows[b].y) (at most 256, one lane per row) and walks the
// segments in ascending column order; its nonzeros of segment s are
// nz[b * nseg + s] = [p0, p1), at most TILE of them. Each lane keeps its
// row's running sum in a register across segments, so no partial sums go
// through memory, and the order of every sum is still ascending columns (bit
// for bit the single-pass result). All blocks of a launch are resident at
// once and walk the segments nearly in step, so the gathered vector's working
// set stays about one segment wide.
template <typename T>
struct SegCsr {
    int nrows;
    int nseg;
    const int* rowptr;   // nseg * (nrows + 1), absolute offsets
    const int* col;
    const T* val;
    int nblk;
    const int2* rows;    // nblk
    const int2* nz;      // nblk * nseg
    static constexpr int first = 1, last = 1;
};

// Row sums of A * v for block b over every segment; on_row(row, sum, pre) as
// in stream_rows. The next segment's index and value loads are issued before
// the current segment's sums are read from LDS.
template <int TILE, typename T, typename V, typename Pre, typename F>
__device__ __forceinline__ void stream_rows(const SegCsr<T>& A, const V* __restrict__ v, T* prod, int b,
                                            Pre&& prefetch, F&& on_row) {
    static_assert(TILE % kThreads == 0, "tile must be a multiple of the workgroup");
    constexpr int PER = TILE / kThreads;
    const int tid = threadIdx.x;
    const int2 rr2 = A.rows[b];
    const int r0 = rr2.x, r1 = rr2.y;
    if (r1 == r0) return;
    const int r = r0 + tid;
    const bool has_row = r < r1;
    const int rr = has_row ? r : r0;
    const auto pre = prefetch(rr);
    const int2* nzb = A.nz + static_cast<size_t>(b) * A.nseg;
    int ci[PER];
    T vi[PER];
    auto load = [&](int seg, int (&c)[PER], T (&w)[PER]) {
        const int2 pz = nzb[seg];
        const int cnt = pz.y - pz.x;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + u * kThreads;
            const bool in = q < cnt;
            const int slot = in ? pz.x + q : 0;
            const int cl = load_stream(A.col + slot);
            const T vl = load_stream(A.val + slot);
            c[u] = in ? cl : 0;
            w[u] = in ? vl : T(0);
        }
    };
    load(0, ci, vi);
    T s = T(0);
    for (int seg = 0; seg < A.nseg; ++seg) {
        const int* rp = A.rowptr + static_cast<size_t>(seg) * (A.nrows + 1);
        const int p0 = nzb[seg].x;
        const int ra = rp[rr] - p0, re = rp[rr + 1] - p0;
        T g[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) g[u] = static_cast<T>(v[ci[u]]);
#pragma unroll
        for (int u = 0; u < PER; ++u) prod[lds_at(tid + u * kThreads)] = vi[u] * g[u];
        __syncthreads();
        if (seg + 1 < A.nseg) load(seg + 1, ci, vi);
        if (has_row) {
            int q = ra;
            for (; q + 4 <= re; q += 4) {
                const T a0 = prod[lds_at(q)], a1 = prod[lds_at(q + 1)], a2 = prod[lds_at(q + 2)], a3 = prod[lds_at(q + 3)];
                s += a0;
                s += a1;
                s += a2;
                s += a3;
            }
            for (; q < re; ++q) s += prod[lds_at(q)];
        }
        __syncthreads();
    }
    if (has_row) on_row(r, s, pre);
}

template <int TILE, typename T, typename V, typename F>
__device__ __forceinline__ void stream_rows(const SegCsr<T>& A, const V* __restrict__ v, T* prod, int b, F&& on_row) {
    stream_rows<TILE>(A, v, prod, b, PreNone{}, [&](int r, T s, NoPre) { on_row(r, s); });
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
