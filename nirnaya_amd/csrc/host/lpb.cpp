// Binary LP container (".nlpb"): the model's compressed-column arrays as they
// sit in memory, so a large LP loads with one read and no text parsing (an MPS
// file of the 400k x 1.2M instance is hundreds of MB of text). It sits beside
// the MPS reader (mps_reader.hpp:15-304) and carries the same model: sense,
// offset, costs, bounds, variable types, matrix, quadratic part and names.
//
// Layout (little-endian; every array starts on a 64-byte boundary, zero
// padding between):
//   header, 128 bytes
//     0  magic "NIRLPB01"
//     8  uint32 header size (128)
//    12  uint32 flags: 1 names, 2 variable types, 4 quadratic part
//    16  int64 rows, 24 int64 columns, 32 int64 nonzeros, 40 int64 Q nonzeros
//    48  int32 sense (1 minimize, -1 maximize), 52 int32 0
//    56  double objective offset
//    64  uint64 payload bytes (file size - 128)
//    72  uint64 payload checksum (below)
//    80  zero to 128
//   payload, in this order
//     int64 start[cols + 1], int32 index[nnz], double value[nnz]
//     double cost[cols], lower[cols], upper[cols], row lower[rows], row upper[rows]
//     int8 type[cols]                                    (flag 2)
//     int64 qstart[cols + 1], int32 qindex[qnnz], double qvalue[qnnz]   (flag 4)
//     names: uint32 length + bytes, for the model, each column, each row  (flag 1)
// Checksum: h = 0xcbf29ce484222325; for each 8-byte little-endian word w of
// the payload h = (h ^ w) * 0x100000001b3 (mod 2^64), then each trailing byte
// the same way.
//
// The reader checks every count against the file size before it touches an
// array, the checksum, and that the matrix is canonical (ascending rows per
// column, no duplicates, no zeros, indices in range) - the form the CSC
// assembly produces (csc_matrix.hpp:30-96) and the engine relies on.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>

#include "prep.hpp"

namespace nir {
namespace {

constexpr char kMagic[8] = {'N', 'I', 'R', 'L', 'P', 'B', '0', '1'};
constexpr uint32_t kHeader = 128;
constexpr uint32_t kNames = 1, kTypes = 2, kQuad = 4;

struct Header {
    char magic[8];
    uint32_t header_size;
    uint32_t flags;
    int64_t nrow, ncol, nnz, qnnz;
    int32_t sense, zero;
    double offset;
    uint64_t payload, checksum;
    uint8_t pad[48];
};
static_assert(sizeof(Header) == kHeader, "header layout");

struct Hash {
    uint64_t h = 0xcbf29ce484222325ull;
    unsigned char tail[8];
    int ntail = 0;
    void word(uint64_t w) { h = (h ^ w) * 0x100000001b3ull; }
    void bytes(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        while (n > 0 && ntail > 0) {   // finish a partial word first
            tail[ntail++] = *b++;
            --n;
            if (ntail == 8) {
                uint64_t w;
                std::memcpy(&w, tail, 8);
                word(w);
                ntail = 0;
            }
        }
        for (; n >= 8; n -= 8, b += 8) {
            uint64_t w;
            std::memcpy(&w, b, 8);
            word(w);
        }
        while (n > 0) {
            tail[ntail++] = *b++;
            --n;
        }
    }
    uint64_t finish() {
        for (int i = 0; i < ntail; ++i) h = (h ^ tail[i]) * 0x100000001b3ull;
        ntail = 0;
        return h;
    }
};

size_t pad64(size_t at) { return (at + 63) & ~static_cast<size_t>(63); }

class Writer {
public:
    explicit Writer(FILE* f) : f_(f) {}
    bool put(const void* p, size_t n) {
        if (n && std::fwrite(p, 1, n, f_) != n) return false;
        hash_.bytes(p, n);
        at_ += n;
        return true;
    }
    bool align() {
        static const unsigned char zeros[64] = {};
        const size_t to = pad64(at_);
        return put(zeros, to - at_);
    }
    template <typename T>
    bool array(const std::vector<T>& v) {
        return align() && put(v.data(), v.size() * sizeof(T));
    }
    bool str(const std::string& s) {
        const uint32_t n = static_cast<uint32_t>(s.size());
        return put(&n, 4) && put(s.data(), s.size());
    }
    size_t size() const { return at_; }
    uint64_t checksum() { return hash_.finish(); }

private:
    FILE* f_;
    Hash hash_;
    size_t at_ = kHeader;   // offsets count from the start of the file
};

// Bounds-checked cursor over the mapped payload.
class Cursor {
public:
    Cursor(const unsigned char* base, size_t size) : base_(base), size_(size) {}
    template <typename T>
    bool array(std::vector<T>& out, int64_t count, std::string& err, const char* what) {
        at_ = pad64(at_);
        if (count < 0 || static_cast<uint64_t>(count) > (size_ - std::min(at_, size_)) / sizeof(T)) {
            err = std::string("truncated ") + what;
            return false;
        }
        out.resize(static_cast<size_t>(count));
        if (count) std::memcpy(out.data(), base_ + at_, static_cast<size_t>(count) * sizeof(T));
        at_ += static_cast<size_t>(count) * sizeof(T);
        return true;
    }
    bool str(std::string& out, std::string& err) {
        uint32_t n;
        if (size_ - std::min(at_, size_) < 4) {
            err = "truncated names";
            return false;
        }
        std::memcpy(&n, base_ + at_, 4);
        at_ += 4;
        if (n > size_ - at_) {
            err = "truncated names";
            return false;
        }
        out.assign(reinterpret_cast<const char*>(base_ + at_), n);
        at_ += n;
        return true;
    }
    void align() { at_ = pad64(at_); }
    size_t at() const { return at_; }

private:
    const unsigned char* base_;
    size_t size_;
    size_t at_ = kHeader;
};

bool canonical(const CscMatrix& A, std::string& err, const char* what) {
    // check() walks start[] into index[]: first make sure start[] cannot
    // point past the arrays.
    if (A.start.size() != static_cast<size_t>(A.ncol) + 1 || A.start.front() != 0 ||
        A.start.back() != static_cast<Off>(A.index.size()) || A.index.size() != A.value.size()) {
        err = std::string(what) + ": column starts do not match the nonzero count";
        return false;
    }
    const std::string e = A.check();
    if (!e.empty()) {
        err = std::string(what) + ": " + e;
        return false;
    }
    for (double v : A.value)
        if (!std::isfinite(v) || v == 0.0) {
            err = std::string(what) + ": zero or non-finite coefficient";
            return false;
        }
    return true;
}

}  // namespace

bool is_lpb_file(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char m[8];
    const bool ok = std::fread(m, 1, 8, f) == 8 && std::memcmp(m, kMagic, 8) == 0;
    std::fclose(f);
    return ok;
}

bool write_lpb(const std::string& path, const Model& m, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot open " + path + " for writing";
        return false;
    }
    Header h{};
    std::memcpy(h.magic, kMagic, 8);
    h.header_size = kHeader;
    const bool names = !m.col_names.empty() || !m.row_names.empty();
    bool types = false;
    for (VarType t : m.vartype) types = types || t != VarType::Continuous;
    const bool quad = m.Q.nnz() > 0;
    h.flags = (names ? kNames : 0) | (types ? kTypes : 0) | (quad ? kQuad : 0);
    h.nrow = m.nrow;
    h.ncol = m.ncol;
    h.nnz = m.A.nnz();
    h.qnnz = quad ? m.Q.nnz() : 0;
    h.sense = static_cast<int32_t>(m.sense);
    h.offset = m.obj_offset;
    bool ok = std::fwrite(&h, 1, kHeader, f) == kHeader;
    Writer w(f);
    ok = ok && w.array(m.A.start) && w.array(m.A.index) && w.array(m.A.value) && w.array(m.cost) &&
         w.array(m.collo) && w.array(m.colup) && w.array(m.rowlo) && w.array(m.rowup);
    if (ok && types) {
        std::vector<int8_t> t(static_cast<size_t>(m.ncol), 0);
        for (size_t j = 0; j < t.size() && j < m.vartype.size(); ++j) t[j] = static_cast<int8_t>(m.vartype[j]);
        ok = w.array(t);
    }
    if (ok && quad) ok = w.array(m.Q.start) && w.array(m.Q.index) && w.array(m.Q.value);
    if (ok && names) {
        ok = w.align() && w.str(m.name);
        for (Idx j = 0; ok && j < m.ncol; ++j) ok = w.str(j < static_cast<Idx>(m.col_names.size()) ? m.col_names[j] : "");
        for (Idx i = 0; ok && i < m.nrow; ++i) ok = w.str(i < static_cast<Idx>(m.row_names.size()) ? m.row_names[i] : "");
    }
    if (ok) {
        h.payload = w.size() - kHeader;
        h.checksum = w.checksum();
        ok = std::fseek(f, 0, SEEK_SET) == 0 && std::fwrite(&h, 1, kHeader, f) == kHeader;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) err = "write error on " + path;
    return ok;
}

Status read_lpb(const std::string& path, Model& out, std::string& err) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) {
        err = "cannot open " + path;
        return Status::InputError;
    }
    struct stat sb;
    if (::fstat(fd, &sb) != 0 || sb.st_size < static_cast<off_t>(kHeader)) {
        ::close(fd);
        err = "not a binary LP file (too short)";
        return Status::InputError;
    }
    const size_t size = static_cast<size_t>(sb.st_size);
    void* map = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (map == MAP_FAILED) {
        err = "cannot map " + path;
        return Status::InputError;
    }
    const unsigned char* base = static_cast<const unsigned char*>(map);
    struct Unmap {
        void* p;
        size_t n;
        ~Unmap() { ::munmap(p, n); }
    } unmap{map, size};

    Header h;
    std::memcpy(&h, base, kHeader);
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.header_size != kHeader) {
        err = "not a binary LP file (bad magic)";
        return Status::InputError;
    }
    const int64_t kMaxIdx = std::numeric_limits<Idx>::max() - 1;
    if (h.nrow < 0 || h.ncol < 0 || h.nrow > kMaxIdx || h.ncol > kMaxIdx || h.nnz < 0 || h.qnnz < 0 ||
        (h.sense != 1 && h.sense != -1) || (h.flags & ~(kNames | kTypes | kQuad)) != 0) {
        err = "corrupt header";
        return Status::InputError;
    }
    if (h.payload != size - kHeader) {
        err = "size mismatch (truncated or extended file)";
        return Status::InputError;
    }
    Hash hs;
    hs.bytes(base + kHeader, size - kHeader);
    if (hs.finish() != h.checksum) {
        err = "checksum mismatch";
        return Status::InputError;
    }

    Model m;
    m.sense = h.sense == 1 ? Sense::Minimize : Sense::Maximize;
    m.obj_offset = h.offset;
    m.nrow = static_cast<Idx>(h.nrow);
    m.ncol = static_cast<Idx>(h.ncol);
    m.A.nrow = m.nrow;
    m.A.ncol = m.ncol;
    Cursor c(base, size);
    bool ok = c.array(m.A.start, h.ncol + 1, err, "column starts") && c.array(m.A.index, h.nnz, err, "row indices") &&
              c.array(m.A.value, h.nnz, err, "coefficients") && c.array(m.cost, h.ncol, err, "costs") &&
              c.array(m.collo, h.ncol, err, "lower bounds") && c.array(m.colup, h.ncol, err, "upper bounds") &&
              c.array(m.rowlo, h.nrow, err, "row lower bounds") && c.array(m.rowup, h.nrow, err, "row upper bounds");
    m.vartype.assign(static_cast<size_t>(m.ncol), VarType::Continuous);
    if (ok && (h.flags & kTypes)) {
        std::vector<int8_t> t;
        ok = c.array(t, h.ncol, err, "variable types");
        for (size_t j = 0; ok && j < t.size(); ++j) {
            if (t[j] < 0 || t[j] > 3) {
                err = "corrupt variable type";
                ok = false;
            } else {
                m.vartype[j] = static_cast<VarType>(t[j]);
            }
        }
    }
    if (ok && (h.flags & kQuad)) {
        m.Q = CscMatrix(m.ncol, m.ncol);
        ok = c.array(m.Q.start, h.ncol + 1, err, "quadratic starts") &&
             c.array(m.Q.index, h.qnnz, err, "quadratic indices") && c.array(m.Q.value, h.qnnz, err, "quadratic values");
    }
    if (ok && (h.flags & kNames)) {
        c.align();   // the names start on a 64-byte boundary
        ok = c.str(m.name, err);
        m.col_names.resize(static_cast<size_t>(m.ncol));
        m.row_names.resize(static_cast<size_t>(m.nrow));
        for (Idx j = 0; ok && j < m.ncol; ++j) ok = c.str(m.col_names[j], err);
        for (Idx i = 0; ok && i < m.nrow; ++i) ok = c.str(m.row_names[i], err);
    }
    if (ok && !(h.flags & kNames)) {
        // Every model carries names (the builder's defaults, model.cpp: "C<j>", "R<i>").
        m.col_names.reserve(static_cast<size_t>(m.ncol));
        m.row_names.reserve(static_cast<size_t>(m.nrow));
        for (Idx j = 0; j < m.ncol; ++j) m.col_names.push_back("C" + std::to_string(j));
        for (Idx i = 0; i < m.nrow; ++i) m.row_names.push_back("R" + std::to_string(i));
    }
    if (ok && c.at() != size) {
        err = "trailing bytes after the last section";
        ok = false;
    }
    if (!ok) return Status::InputError;
    if (!canonical(m.A, err, "matrix") || ((h.flags & kQuad) && !canonical(m.Q, err, "quadratic"))) {
        return Status::InputError;
    }
    out = std::move(m);
    return Status::Optimal;
}

Status read_model(const std::string& path, Model& out, std::string& err) {
    return is_lpb_file(path) ? read_lpb(path, out, err) : read_mps(path, out, err);
}

}  // namespace nir
