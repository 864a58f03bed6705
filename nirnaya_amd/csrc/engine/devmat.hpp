// Device-resident sparse matrices for the CSR-stream kernels: HIP error
// checking, owning device arrays, the row-block partition and the
// column-segmented layout (see spmv.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <stdexcept>
#include <string>
#include <vector>

#include "spmv.hpp"

namespace nir {
namespace gpu {

struct HipError : std::runtime_error {
    explicit HipError(const std::string& what) : std::runtime_error(what) {}
};

#define HIP_CHECK(expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::nir::gpu::HipError(std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    ~DevArray() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        n = count;
        HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T)));
    }
    void upload(const T* src, size_t count, hipStream_t s) {
        if (count > n) throw std::runtime_error("upload overflow");
        if (count) HIP_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* dst, size_t count, hipStream_t s) const {
        if (count > n) throw std::runtime_error("download overflow");
        if (count) HIP_CHECK(hipMemcpyAsync(dst, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// Row blocks for CSR-stream: <= tile_rows rows and <= tile_nnz nonzeros per
// block, or one row longer than tile_nnz on its own. Always >= 1 block.
inline std::vector<int> partition_rows(const int* rowptr, int nrows, int tile_nnz, int tile_rows) {
    std::vector<int> blk{0};
    int r = 0;
    while (r < nrows) {
        const int s = r;
        if (rowptr[r + 1] - rowptr[r] > tile_nnz) {
            ++r;
        } else {
            while (r < nrows && r - s < tile_rows && rowptr[r + 1] - rowptr[s] <= tile_nnz) ++r;
        }
        blk.push_back(r);
    }
    if (blk.size() == 1) blk.push_back(0);
    return blk;
}

// Tile shape of one matrix: LDS tile (values), rows per workgroup and the
// column-segment width (0 = one segment).
struct TileShape {
    int tile = 2048;
    int rows = 256;
    int seg_width = 0;
    int fused = 0;   // walk all segments in one launch (SegCsr) when the rows allow it
    int split = 0;   // split rows of more than kLongRow nonzeros over workgroups (one segment
                     // only; for kernels that sum such rows in the chunked order)
};

// Structure (row pointers per segment, column indices, row blocks) of a
// segmented CSR matrix; value arrays of either precision are kept beside it.
struct CsrShape {
    int nrows = 0, ncols = 0, nnz = 0, nseg = 1;
    TileShape shape;
    DevArray<int> rowptr;   // nseg * (nrows + 1), absolute offsets
    DevArray<int> col;      // nnz, segment-major
    DevArray<int4> desc;    // concatenated per-segment block descriptors
    std::vector<int> blk_off, nblk;
    std::vector<long long> seg_nz;   // nseg + 1: segment s holds slots [seg_nz[s], seg_nz[s + 1])
    // Fused-segment blocks (SegCsr): row ranges and per-segment nonzero ranges.
    DevArray<int2> frows, fnz;
    int fnblk = 0;
    bool fused_ok = false;
    // Rows split over workgroups (TileShape::split).
    DevArray<int4> lrow;     // per split row {first group slot, groups, pieces, row}
    DevArray<int> lcount;    // per split row: zero between launches
    DevArray<double> gsum;   // group sums, FP64 or FP32 values
    int nsplit = 0, ngroups = 0, npieces = 0;
    // Per segment: how many blocks hold one row longer than a tile or a piece
    // of a split row (kernels compiled without the long-row code need none).
    std::vector<int> nlong;

    template <typename T>
    Csr<T> view(const T* val, int s) const {
        return Csr<T>{nrows, ncols, rowptr.p + static_cast<size_t>(s) * (nrows + 1), col.p, val, nblk[s],
                      desc.p + blk_off[s], lrow.p, lcount.p, reinterpret_cast<T*>(gsum.p)};
    }
    template <typename T>
    Pass<T> pass(const T* val, T* part, int s) const {
        return Pass<T>{view(val, s), part, s == 0 ? 1 : 0, s == nseg - 1 ? 1 : 0};
    }
    template <typename T>
    SegCsr<T> fused(const T* val) const {
        return SegCsr<T>{nrows, nseg, rowptr.p, col.p, val, fnblk, frows.p, fnz.p};
    }
    // Blocks whose per-block results (e.g. violation maxima) the last launch writes.
    int result_blocks() const { return fused_ok ? fnblk : nblk[nseg - 1]; }
};

// Row length above which the layout's rows are split (none: INT_MAX).
inline int split_len(const CsrShape& S) { return S.nsplit > 0 ? kLongRow : INT_MAX; }

// Fused-segment partition: consecutive rows, at most `rows_max`, whose
// nonzeros in every segment fit one tile. Fails (returns false) when a single
// row has more than `tile` nonzeros in some segment.
inline bool partition_fused(const std::vector<int>& rp, int nrows, int nseg, int tile, int rows_max,
                            std::vector<int2>& rows, std::vector<int2>& nz) {
    rows.clear();
    nz.clear();
    auto seg_ptr = [&](int s, int r) { return rp[static_cast<size_t>(s) * (nrows + 1) + r]; };
    int r = 0;
    while (r < nrows) {
        const int s0 = r;
        while (r < nrows && r - s0 < rows_max) {
            bool fits = true;
            for (int s = 0; s < nseg && fits; ++s) fits = seg_ptr(s, r + 1) - seg_ptr(s, s0) <= tile;
            if (!fits) break;
            ++r;
        }
        if (r == s0) return false;
        rows.push_back(make_int2(s0, r));
        for (int s = 0; s < nseg; ++s) nz.push_back(make_int2(seg_ptr(s, s0), seg_ptr(s, r)));
    }
    if (rows.empty()) {   // empty matrix: one block without rows
        rows.push_back(make_int2(0, 0));
        for (int s = 0; s < nseg; ++s) nz.push_back(make_int2(0, 0));
    }
    return true;
}

// Host side of a column-segmented layout of a CSR matrix (columns ascending
// within rows): per segment row pointers (absolute offsets), the stored
// slots' column indices, and `perm`, the input nonzero each stored slot holds.
struct SegHost {
    int nrows = 0, ncols = 0, nnz = 0, nseg = 1;
    std::vector<int> rp;              // nseg * (nrows + 1)
    std::vector<int> perm;            // nnz
    std::vector<int> col;             // nnz, segment-major
    std::vector<long long> seg_nz;    // nseg + 1
};

inline SegHost build_segments(int nrows, int ncols, const std::vector<int>& rowptr, const int* col, int seg_width) {
    SegHost H;
    H.nrows = nrows;
    H.ncols = ncols;
    H.nnz = rowptr.back();
    const int W = (seg_width > 0 && seg_width < ncols) ? seg_width : std::max(ncols, 1);
    H.nseg = std::max(1, (ncols + W - 1) / W);
    const int nseg = H.nseg;
    H.rp.assign(static_cast<size_t>(nseg) * (nrows + 1), 0);
    H.perm.assign(static_cast<size_t>(H.nnz), 0);
    // Count entries per segment, then lay segments out one after another.
    std::vector<long long> seg_count(nseg, 0);
    for (int p = 0; p < H.nnz; ++p) ++seg_count[col[p] / W];
    H.seg_nz.assign(nseg + 1, 0);
    for (int s = 0; s < nseg; ++s) H.seg_nz[s + 1] = H.seg_nz[s] + seg_count[s];
    std::vector<long long> fill(H.seg_nz.begin(), H.seg_nz.end() - 1);
    for (int s = 0; s < nseg; ++s) H.rp[static_cast<size_t>(s) * (nrows + 1)] = static_cast<int>(H.seg_nz[s]);
    for (int r = 0; r < nrows; ++r) {
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            const int s = col[p] / W;
            H.perm[fill[s]++] = p;
        }
        for (int s = 0; s < nseg; ++s) H.rp[static_cast<size_t>(s) * (nrows + 1) + r + 1] = static_cast<int>(fill[s]);
    }
    H.col.resize(static_cast<size_t>(H.nnz));
    for (int q = 0; q < H.nnz; ++q) H.col[q] = col[H.perm[q]];
    return H;
}

// Row pointers and column indices of a layout on the device.
inline void upload_structure(CsrShape& S, const SegHost& H, hipStream_t st) {
    S.nrows = H.nrows;
    S.ncols = H.ncols;
    S.nnz = H.nnz;
    S.nseg = H.nseg;
    S.seg_nz = H.seg_nz;
    S.rowptr.alloc(H.rp.size());
    S.rowptr.upload(H.rp.data(), H.rp.size(), st);
    S.col.alloc(H.col.size());
    S.col.upload(H.col.data(), H.col.size(), st);
}

// Pieces of a split row: whole groups of kChunk * kGroupChunks nonzeros,
// max(tile, one group) nonzeros each, from the row's first nonzero.
inline int split_piece_len(int tile) { return std::max(tile, kGroupLen); }

// Block descriptors of one segment's row pointers `srp` (appended to desc),
// with rows of more than kLongRow nonzeros split into pieces when `split`
// (their {first group slot, groups, pieces, row} appended to lrow,
// group slots counted in `groups`).
inline void segment_blocks(const int* srp, int nrows, const TileShape& shape, bool split, std::vector<int4>& desc,
                           std::vector<int4>& lrow, int& groups) {
    const int piece_len = split_piece_len(shape.tile);
    if (split && piece_len > 2 * kGroupLen) throw std::invalid_argument("split rows: tile above 4096 nonzeros");
    const std::vector<int> b = partition_rows(srp, nrows, shape.tile, std::min(shape.rows, kThreads));
    for (size_t k = 0; k + 1 < b.size(); ++k) {
        const int p0 = srp[b[k]], p1 = srp[b[k + 1]];
        if (split && b[k + 1] == b[k] + 1 && p1 - p0 > kLongRow) {
            const int id = static_cast<int>(lrow.size());
            const int ng = (p1 - p0 + kGroupLen - 1) / kGroupLen;
            const int np = (p1 - p0 + piece_len - 1) / piece_len;
            lrow.push_back(make_int4(groups, ng, np, b[k]));
            groups += ng;
            for (int q = p0; q < p1; q += piece_len) desc.push_back(make_int4(b[k], -(id + 1), q, std::min(p1, q + piece_len)));
            continue;
        }
        desc.push_back(make_int4(b[k], b[k + 1], p0, p1));
    }
}

// Row blocks of a layout for one tile shape (per-segment descriptors, and the
// fused-segment partition when asked for and possible), on the device.
inline void set_blocks(CsrShape& S, const SegHost& H, TileShape shape, hipStream_t st) {
    S.shape = shape;
    const int nseg = H.nseg, nrows = H.nrows;
    std::vector<int4> desc_all, lrow;
    int groups = 0;
    S.blk_off.assign(nseg, 0);
    S.nblk.assign(nseg, 0);
    for (int s = 0; s < nseg; ++s) {
        S.blk_off[s] = static_cast<int>(desc_all.size());
        segment_blocks(H.rp.data() + static_cast<size_t>(s) * (nrows + 1), nrows, shape, shape.split && nseg == 1,
                       desc_all, lrow, groups);
        S.nblk[s] = static_cast<int>(desc_all.size()) - S.blk_off[s];
    }
    S.desc.alloc(desc_all.size());
    S.desc.upload(desc_all.data(), desc_all.size(), st);
    S.nlong.assign(nseg, 0);
    for (int s = 0; s < nseg; ++s)
        for (int k = 0; k < S.nblk[s]; ++k) {
            const int4 d = desc_all[static_cast<size_t>(S.blk_off[s]) + k];
            if (d.y < 0 || (d.y == d.x + 1 && d.w - d.z > shape.tile)) ++S.nlong[s];
        }
    S.nsplit = static_cast<int>(lrow.size());
    S.ngroups = groups;
    S.npieces = 0;
    for (const int4& L : lrow) S.npieces += L.z;
    S.lrow.release();
    S.lcount.release();
    S.gsum.release();
    if (S.nsplit > 0) {
        S.lrow.alloc(lrow.size());
        S.lrow.upload(lrow.data(), lrow.size(), st);
        S.lcount.alloc(lrow.size());
        HIP_CHECK(hipMemsetAsync(S.lcount.p, 0, lrow.size() * sizeof(int), st));
        S.gsum.alloc(static_cast<size_t>(groups));
    }
    S.fused_ok = false;
    S.fnblk = 0;
    S.frows.release();
    S.fnz.release();
    if (shape.fused && nseg > 1) {
        std::vector<int2> fr, fz;
        if (partition_fused(H.rp, nrows, nseg, shape.tile, std::min(shape.rows, kThreads), fr, fz)) {
            S.fused_ok = true;
            S.fnblk = static_cast<int>(fr.size());
            S.frows.alloc(fr.size());
            S.frows.upload(fr.data(), fr.size(), st);
            S.fnz.alloc(fz.size());
            S.fnz.upload(fz.data(), fz.size(), st);
        }
    }
}

// Upload a CSR matrix (columns ascending within rows) in the segmented
// layout. `perm` receives, for each stored slot, the input nonzero it holds.
inline void upload_shape(CsrShape& S, int nrows, int ncols, const std::vector<int>& rowptr, const int* col,
                         TileShape shape, std::vector<int>& perm, hipStream_t st) {
    SegHost H = build_segments(nrows, ncols, rowptr, col, shape.seg_width);
    upload_structure(S, H, st);
    set_blocks(S, H, shape, st);
    HIP_CHECK(hipStreamSynchronize(st));
    perm.swap(H.perm);
}

// A matrix on the device: segmented structure plus its FP64 values and the
// running-sum buffer of the segment passes.
struct DevMatrix {
    CsrShape shape;
    DevArray<double> val;
    DevArray<double> part;
};

inline void upload_matrix(DevMatrix& M, int nrows, int ncols, const std::vector<int>& rowptr, const int* col,
                          const double* val, TileShape ts, hipStream_t s) {
    std::vector<int> perm;
    upload_shape(M.shape, nrows, ncols, rowptr, col, ts, perm, s);
    std::vector<double> v(perm.size());
    for (size_t q = 0; q < perm.size(); ++q) v[q] = val[perm[q]];
    M.val.alloc(v.size());
    M.val.upload(v.data(), v.size(), s);
    M.part.alloc(static_cast<size_t>(std::max(nrows, 1)));
    HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace gpu
}  // namespace nir

// Instantiate a tile-templated launch for the matrix's LDS tile size.
#define TILE_DISPATCH(tile, ...)     \
    switch (tile) {                  \
        case 1024: {                 \
            constexpr int TL = 1024; \
            __VA_ARGS__;             \
        } break;                     \
        case 4096: {                 \
            constexpr int TL = 4096; \
            __VA_ARGS__;             \
        } break;                     \
        default: {                   \
            constexpr int TL = 2048; \
            __VA_ARGS__;             \
        } break;                     \
    }
