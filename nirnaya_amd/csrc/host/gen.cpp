// Seeded synthetic LP families.
//
// RNG: xoshiro256** (Blackman & Vigna) seeded through splitmix64; normals by
// Box-Muller, Poisson by Knuth's product method, k distinct rows by Floyd's
// sampling followed by a sort. Right-hand sides are A x0 for a known x0
// (plus nonnegative slack on inequality rows) and costs are A^T y0 + s with
// s >= 0 and y0 sign-feasible, so every instance is primal and dual feasible.
#include <algorithm>
#include <cstring>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../../include/nirnaya/pdhg.h"
#include "model.hpp"
#include "prep.hpp"

namespace nir {
// The engine's per-thread error text, read by nir_last_error() (engine.hip).
void record_last_error(const std::string& what);
}  // namespace nir

namespace nir {
namespace {

class Rng {
public:
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s_) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ULL;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebULL;
            w = t ^ (t >> 31);
        }
    }
    uint64_t next() {
        const uint64_t out = rotl(s_[1] * 5, 7) * 9;
        const uint64_t t = s_[1] << 17;
        s_[2] ^= s_[0];
        s_[3] ^= s_[1];
        s_[1] ^= s_[2];
        s_[0] ^= s_[3];
        s_[2] ^= t;
        s_[3] = rotl(s_[3], 45);
        return out;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }   // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    uint64_t below(uint64_t n) {   // unbiased integer in [0, n)
        const uint64_t lim = (~uint64_t(0)) - (~uint64_t(0)) % n;
        uint64_t v;
        do v = next(); while (v >= lim);
        return v % n;
    }
    double normal() {
        if (have_spare_) {
            have_spare_ = false;
            return spare_;
        }
        double u1;
        do u1 = uniform(); while (u1 <= 0.0);
        const double u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        spare_ = r * std::sin(2.0 * M_PI * u2);
        have_spare_ = true;
        return r * std::cos(2.0 * M_PI * u2);
    }
    int poisson(double mean) {
        const double limit = std::exp(-mean);
        int k = 0;
        double p = 1.0;
        do {
            ++k;
            p *= uniform();
        } while (p > limit);
        return k - 1;
    }
    // k distinct values from [0, n), ascending (Floyd).
    void distinct(int32_t n, int32_t k, std::vector<int32_t>& out) {
        out.clear();
        if (k >= n) {
            for (int32_t i = 0; i < n; ++i) out.push_back(i);
            return;
        }
        std::unordered_set<int32_t> seen;
        for (int32_t j = n - k; j < n; ++j) {
            const int32_t t = static_cast<int32_t>(below(static_cast<uint64_t>(j) + 1));
            if (seen.insert(t).second) out.push_back(t);
            else { seen.insert(j); out.push_back(j); }
        }
        std::sort(out.begin(), out.end());
    }

private:
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t s_[4];
    bool have_spare_ = false;
    double spare_ = 0.0;
};

}  // namespace
}  // namespace nir

struct nir_owned_lp {
    int32_t m = 0, n = 0;
    double obj_offset = 0.0;
    std::vector<int64_t> start;
    std::vector<int32_t> index;
    std::vector<double> value, cost, collo, colup, rowlo, rowup, x0, y0;   // x0, y0 empty when unknown
};

namespace {

using nir::Rng;

// b = A x0 accumulated column by column; g = A^T y0 one column at a time.
void activities(const nir_owned_lp& g, std::vector<double>& ax) {
    ax.assign(static_cast<size_t>(g.m), 0.0);
    for (int32_t j = 0; j < g.n; ++j)
        for (int64_t p = g.start[j]; p < g.start[j + 1]; ++p) ax[g.index[p]] += g.value[p] * g.x0[j];
}

void costs_from_dual(nir_owned_lp& g, Rng& rng) {
    g.cost.resize(static_cast<size_t>(g.n));
    for (int32_t j = 0; j < g.n; ++j) {
        double acc = 0.0;
        for (int64_t p = g.start[j]; p < g.start[j + 1]; ++p) acc += g.value[p] * g.y0[g.index[p]];
        g.cost[j] = acc + rng.uniform();
    }
}

}  // namespace

extern "C" {

NIR_API nir_owned_lp* nir_gen_equality_lp(uint64_t seed, int32_t m, int32_t n, int32_t per_col) {
    if (m <= 0 || n <= 0 || per_col <= 0) return nullptr;
    auto* g = new nir_owned_lp;
    g->m = m;
    g->n = n;
    Rng rng(seed);
    std::vector<int32_t> rows;
    g->start.assign(1, 0);
    for (int32_t j = 0; j < n; ++j) {
        rng.distinct(m, per_col, rows);
        for (int32_t r : rows) {
            g->index.push_back(r);
            g->value.push_back(rng.normal());
        }
        g->start.push_back(static_cast<int64_t>(g->index.size()));
    }
    g->x0.resize(static_cast<size_t>(n));
    for (auto& v : g->x0) v = rng.uniform();
    std::vector<double> ax;
    activities(*g, ax);
    g->rowlo = ax;
    g->rowup = ax;
    g->collo.assign(static_cast<size_t>(n), 0.0);
    g->colup.assign(static_cast<size_t>(n), 10.0);
    g->y0.resize(static_cast<size_t>(m));
    for (auto& v : g->y0) v = rng.normal();
    costs_from_dual(*g, rng);
    return g;
}

NIR_API nir_owned_lp* nir_gen_mixed_lp(uint64_t seed, int32_t m, int32_t n, double mean, double bad_frac,
                                       double bad_scale) {
    if (m <= 1 || n <= 0) return nullptr;
    auto* g = new nir_owned_lp;
    g->m = m;
    g->n = n;
    Rng rng(seed);
    std::vector<int32_t> rows;
    g->start.assign(1, 0);
    for (int32_t j = 0; j < n; ++j) {
        const int32_t deg = std::min<int32_t>(m, 2 + rng.poisson(mean));
        rng.distinct(m, deg, rows);
        for (int32_t r : rows) {
            g->index.push_back(r);
            g->value.push_back(rng.normal());
        }
        g->start.push_back(static_cast<int64_t>(g->index.size()));
    }
    const int32_t nbad = static_cast<int32_t>(std::llround(bad_frac * n));
    std::vector<int32_t> bad;
    rng.distinct(n, nbad, bad);
    for (int32_t j : bad)
        for (int64_t p = g->start[j]; p < g->start[j + 1]; ++p) g->value[p] *= bad_scale;
    g->collo.assign(static_cast<size_t>(n), 0.0);
    g->colup.resize(static_cast<size_t>(n));
    for (auto& u : g->colup) u = rng.uniform(1.0, 100.0);
    g->x0.resize(static_cast<size_t>(n));
    for (auto& v : g->x0) v = rng.uniform();
    std::vector<double> ax;
    activities(*g, ax);
    const int32_t meq = m / 2;
    g->rowlo.resize(static_cast<size_t>(m));
    g->rowup.resize(static_cast<size_t>(m));
    g->y0.resize(static_cast<size_t>(m));
    for (int32_t i = 0; i < m; ++i) {
        if (i < meq) {
            g->rowlo[i] = g->rowup[i] = ax[i];
            g->y0[i] = rng.normal();
        } else {
            g->rowlo[i] = -nir::kInf;
            g->rowup[i] = ax[i] + rng.uniform();
            g->y0[i] = -std::fabs(rng.normal());
        }
    }
    costs_from_dual(*g, rng);
    return g;
}

// Per (period t, product k): production p, inventory s and Q component
// purchases r_q. Rows: inventory balance s_{t-1} + p - s_t = demand
// (equality, links consecutive periods), a yield row and S-1 spec rows over
// (p, r_q) (<=), and per period one capacity row over all products' p (the
// longest rows). Coefficients U[0.1,2] with 10% sign flips.
NIR_API void nir_gen_planning_dims(uint64_t seed, int32_t* T, int32_t* K) {
    // A stream of its own (seed mixed with a constant) so the dimensions do
    // not reuse the instance's coefficient stream.
    Rng rng(seed ^ 0x5bd1e9955bd1e995ULL);
    const int32_t t = 4 + static_cast<int32_t>(rng.below(45));   // U{4..48}
    const int32_t k = 5 + static_cast<int32_t>(rng.below(36));   // U{5..40}
    if (T) *T = t;
    if (K) *K = k;
}

NIR_API nir_owned_lp* nir_gen_staircase_lp(uint64_t seed, int32_t T, int32_t K) {
    if (T <= 0 || K <= 0) return nullptr;
    constexpr int Q = 4;       // components per product
    constexpr int S = 9;       // yield + spec rows per (t, k)
    auto* g = new nir_owned_lp;
    Rng rng(seed);
    const int32_t per = 2 + Q;   // variables per (t, k): p, s, r_0..r_{Q-1}
    g->n = T * K * per;
    g->m = T * K * (1 + S) + T;
    auto var = [&](int t, int k, int which) { return (t * K + k) * per + which; };   // which: 0=p 1=s 2+q=r_q
    auto coef = [&]() {
        const double a = rng.uniform(0.1, 2.0);
        return rng.uniform() < 0.1 ? -a : a;
    };
    std::vector<nir::Entry> ent;
    std::vector<char> is_eq(static_cast<size_t>(g->m), 0);
    std::vector<double> demand(static_cast<size_t>(g->m), 0.0);
    int32_t row = 0;
    for (int t = 0; t < T; ++t) {
        for (int k = 0; k < K; ++k) {
            // balance: s_{t-1,k} + p_{t,k} - s_{t,k} = d_{t,k}
            if (t > 0) ent.push_back({row, var(t - 1, k, 1), 1.0});
            ent.push_back({row, var(t, k, 0), 1.0});
            ent.push_back({row, var(t, k, 1), -1.0});
            is_eq[row] = 1;
            demand[row] = rng.uniform(10.0, 1000.0);
            ++row;
            // yield and spec rows over p and the components
            for (int s = 0; s < S; ++s) {
                ent.push_back({row, var(t, k, 0), s == 0 ? 1.0 : coef()});
                for (int q = 0; q < Q; ++q) ent.push_back({row, var(t, k, 2 + q), s == 0 ? -std::fabs(coef()) : coef()});
                ++row;
            }
        }
        for (int k = 0; k < K; ++k) ent.push_back({row, var(t, k, 0), std::fabs(coef())});
        ++row;
    }
    // A feasible point: small inventories, production meets demand exactly.
    g->x0.assign(static_cast<size_t>(g->n), 0.0);
    std::vector<double> inv(static_cast<size_t>(T * K));
    for (auto& v : inv) v = rng.uniform(0.0, 5.0);
    row = 0;
    for (int t = 0; t < T; ++t) {
        for (int k = 0; k < K; ++k) {
            const double prev = t > 0 ? inv[(t - 1) * K + k] : 0.0;
            g->x0[var(t, k, 1)] = inv[t * K + k];
            g->x0[var(t, k, 0)] = demand[row] + inv[t * K + k] - prev;
            for (int q = 0; q < Q; ++q) g->x0[var(t, k, 2 + q)] = rng.uniform(1.0, 100.0);
            row += 1 + S;
        }
        row += 1;
    }
    nir::CscMatrix A = nir::CscMatrix::assemble(g->m, g->n, ent);
    g->start = A.start;
    g->index = A.index;
    g->value = A.value;
    std::vector<double> ax;
    activities(*g, ax);
    g->rowlo.resize(static_cast<size_t>(g->m));
    g->rowup.resize(static_cast<size_t>(g->m));
    g->y0.resize(static_cast<size_t>(g->m));
    for (int32_t i = 0; i < g->m; ++i) {
        if (is_eq[i]) {
            g->rowlo[i] = g->rowup[i] = ax[i];
            g->y0[i] = rng.normal();
        } else {
            g->rowlo[i] = -nir::kInf;
            g->rowup[i] = ax[i] + rng.uniform();
            g->y0[i] = -std::fabs(rng.normal());
        }
    }
    g->collo.assign(static_cast<size_t>(g->n), 0.0);
    g->colup.resize(static_cast<size_t>(g->n));
    for (int32_t j = 0; j < g->n; ++j) g->colup[j] = std::max(2.0 * g->x0[j], 1.0) + rng.uniform(0.0, 100.0);
    costs_from_dual(*g, rng);
    return g;
}

NIR_API int nir_owned_lp_view(const nir_owned_lp* g, nir_lp* out) {
    if (!g || !out) return -1;
    out->nrow = g->m;
    out->ncol = g->n;
    out->col_start = g->start.data();
    out->row_index = g->index.data();
    out->value = g->value.data();
    out->cost = g->cost.data();
    out->collo = g->collo.data();
    out->colup = g->colup.data();
    out->rowlo = g->rowlo.data();
    out->rowup = g->rowup.data();
    out->obj_offset = g->obj_offset;
    return 0;
}

NIR_API int64_t nir_owned_lp_nnz(const nir_owned_lp* g) { return g ? g->start.back() : -1; }
NIR_API const double* nir_owned_lp_x0(const nir_owned_lp* g) {
    return (g && !g->x0.empty()) ? g->x0.data() : nullptr;
}
NIR_API const double* nir_owned_lp_y0(const nir_owned_lp* g) {
    return (g && !g->y0.empty()) ? g->y0.data() : nullptr;
}

NIR_API int nir_lp_write_lpb(const nir_lp* lp, const char* path) {
    try {
        if (!lp || !path) throw std::invalid_argument("missing LP or path");
        if (lp->nrow < 0 || lp->ncol < 0) throw std::invalid_argument("negative dimensions");
        nir::Model m;
        m.nrow = lp->nrow;
        m.ncol = lp->ncol;
        const size_t n = static_cast<size_t>(lp->ncol), mr = static_cast<size_t>(lp->nrow);
        if ((n && (!lp->col_start || !lp->cost || !lp->collo || !lp->colup)) || (mr && (!lp->rowlo || !lp->rowup)))
            throw std::invalid_argument("missing LP array");
        m.A.nrow = lp->nrow;
        m.A.ncol = lp->ncol;
        m.A.start.assign(lp->col_start ? lp->col_start : nullptr, lp->col_start ? lp->col_start + n + 1 : nullptr);
        if (m.A.start.empty()) m.A.start.assign(1, 0);
        const size_t nnz = static_cast<size_t>(m.A.start.back());
        if (nnz && (!lp->row_index || !lp->value)) throw std::invalid_argument("missing matrix arrays");
        m.A.index.assign(lp->row_index, lp->row_index + nnz);
        m.A.value.assign(lp->value, lp->value + nnz);
        const std::string defect = m.A.start.front() == 0 ? m.A.check() : std::string("start[0] != 0");
        if (!defect.empty()) throw std::invalid_argument("matrix not canonical: " + defect);
        m.cost.assign(lp->cost, lp->cost + n);
        m.collo.assign(lp->collo, lp->collo + n);
        m.colup.assign(lp->colup, lp->colup + n);
        m.rowlo.assign(lp->rowlo, lp->rowlo + mr);
        m.rowup.assign(lp->rowup, lp->rowup + mr);
        m.vartype.assign(n, nir::VarType::Continuous);
        m.obj_offset = lp->obj_offset;
        std::string err;
        if (!nir::write_lpb(path, m, err)) throw std::runtime_error(err);
        return 0;
    } catch (const std::exception& ex) {
        nir::record_last_error(ex.what());
        return -1;
    }
}

NIR_API nir_owned_lp* nir_lp_read_lpb(const char* path) {
    try {
        if (!path) throw std::invalid_argument("missing path");
        nir::Model m;
        std::string err;
        if (nir::read_lpb(path, m, err) != nir::Status::Optimal) throw std::runtime_error(err);
        if (m.is_qp()) throw std::invalid_argument("the file holds a quadratic objective");
        if (m.is_mip()) throw std::invalid_argument("the file holds integer variables");
        auto g = std::make_unique<nir_owned_lp>();
        g->m = m.nrow;
        g->n = m.ncol;
        // min form: a maximisation is stored as min -c^T x - offset (src/model/model.hpp:73-110)
        const double sgn = m.sense == nir::Sense::Maximize ? -1.0 : 1.0;
        g->obj_offset = sgn * m.obj_offset;
        g->start = std::move(m.A.start);
        g->index = std::move(m.A.index);
        g->value = std::move(m.A.value);
        g->cost = std::move(m.cost);
        for (double& c : g->cost) c *= sgn;
        g->collo = std::move(m.collo);
        g->colup = std::move(m.colup);
        g->rowlo = std::move(m.rowlo);
        g->rowup = std::move(m.rowup);
        return g.release();
    } catch (const std::exception& ex) {
        nir::record_last_error(ex.what());
        return nullptr;
    }
}
NIR_API void nir_owned_lp_free(nir_owned_lp* g) { delete g; }

}  // extern "C"

// ------------------------------------------------------------------ presolve C ABI

extern "C" {

NIR_API int nir_lp_presolve(const nir_lp* lp, const int8_t* integer, int32_t level, nir_presolve_info* info,
                            double* fixed_x, int32_t* col_map, int32_t* row_map, double* collo, double* colup,
                            double* rowlo, double* rowup) {
    try {
        if (!lp || lp->nrow < 0 || lp->ncol < 0) throw std::invalid_argument("invalid LP");
        if (lp->ncol > 0 && (!lp->col_start || !lp->cost || !lp->collo || !lp->colup))
            throw std::invalid_argument("missing column arrays");
        if (lp->nrow > 0 && (!lp->rowlo || !lp->rowup)) throw std::invalid_argument("missing row arrays");
        nir::Model M;
        M.name = "lp";
        M.ncol = lp->ncol;
        M.nrow = lp->nrow;
        const long long nnz = lp->ncol > 0 ? lp->col_start[lp->ncol] : 0;
        M.cost.assign(lp->cost, lp->cost + lp->ncol);
        M.collo.assign(lp->collo, lp->collo + lp->ncol);
        M.colup.assign(lp->colup, lp->colup + lp->ncol);
        M.vartype.assign(static_cast<size_t>(lp->ncol), nir::VarType::Continuous);
        if (integer)
            for (int j = 0; j < lp->ncol; ++j)
                if (integer[j]) M.vartype[j] = nir::VarType::Integer;
        M.rowlo.assign(lp->rowlo, lp->rowlo + lp->nrow);
        M.rowup.assign(lp->rowup, lp->rowup + lp->nrow);
        M.obj_offset = lp->obj_offset;
        M.A = nir::CscMatrix(lp->nrow, lp->ncol);
        if (lp->ncol > 0) {
            M.A.start.assign(lp->col_start, lp->col_start + lp->ncol + 1);
            M.A.index.assign(lp->row_index, lp->row_index + nnz);
            M.A.value.assign(lp->value, lp->value + nnz);
        }
        M.Q = nir::CscMatrix(lp->ncol, lp->ncol);
        for (int j = 0; j < lp->ncol; ++j) M.col_names.push_back("C" + std::to_string(j));
        for (int i = 0; i < lp->nrow; ++i) M.row_names.push_back("R" + std::to_string(i));
        const std::string defect = M.A.check();
        if (!defect.empty()) throw std::invalid_argument("matrix: " + defect);
        const nir::Presolved P = nir::presolve(M, level, 20);
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->status = static_cast<int32_t>(P.status);
            info->solved = P.solved ? 1 : 0;
            info->nrow = P.reduced.nrow;
            info->ncol = P.reduced.ncol;
            info->rows_removed = P.rows_removed;
            info->cols_removed = P.cols_removed;
            info->tightenings = P.tightenings;
            info->obj_offset = P.reduced.obj_offset;
        }
        if (fixed_x) std::copy(P.fixed_x.begin(), P.fixed_x.end(), fixed_x);
        if (col_map) std::copy(P.col_map.begin(), P.col_map.end(), col_map);
        if (row_map) std::copy(P.row_map.begin(), P.row_map.end(), row_map);
        if (collo) std::copy(P.reduced.collo.begin(), P.reduced.collo.end(), collo);
        if (colup) std::copy(P.reduced.colup.begin(), P.reduced.colup.end(), colup);
        if (rowlo) std::copy(P.reduced.rowlo.begin(), P.reduced.rowlo.end(), rowlo);
        if (rowup) std::copy(P.reduced.rowup.begin(), P.reduced.rowup.end(), rowup);
        return 0;
    } catch (const std::exception& ex) {
        nir::record_last_error(ex.what());
        return -1;
    }
}

}  // extern "C"
