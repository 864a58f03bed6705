// ORACLE - test infrastructure only.
//
// Thin extern "C" shims over the reference's own header-only sources, built by
// oracle/Makefile straight from /root/reference (nothing copied) into
// oracle/_ref/libref.so together with the reference's src/driver/c_api.cpp,
// whose sov_* entry points come along unchanged. Used to pin the C oracle and
// to compare whole solves against the reference facade.
#include <cstdint>
#include <cstring>
#include <vector>

#include "sov/solver.hpp"

namespace {

sov::Model make_model(int32_t m, int32_t n, const int64_t* start, const int32_t* index, const double* value,
                      const double* c, const double* l, const double* u, const double* rl, const double* ru) {
    sov::Model M;
    M.name = "harness";
    M.ncol = n;
    M.nrow = m;
    M.cost.assign(c, c + n);
    M.collo.assign(l, l + n);
    M.colup.assign(u, u + n);
    M.vartype.assign(static_cast<size_t>(n), sov::VarType::Continuous);
    M.rowlo.assign(rl, rl + m);
    M.rowup.assign(ru, ru + m);
    M.A = sov::CscMatrix(m, n);
    M.A.start.assign(start, start + n + 1);
    M.A.index.assign(index, index + start[n]);
    M.A.value.assign(value, value + start[n]);
    M.Q = sov::CscMatrix(n, n);
    for (int32_t j = 0; j < n; ++j) M.col_names.push_back("C" + std::to_string(j));
    for (int32_t i = 0; i < m; ++i) M.row_names.push_back("R" + std::to_string(i));
    return M;
}

}  // namespace

extern "C" {

// PdhgSolver::solve on the LP as given. simplex_iter_limit sets the cap
// max(5000, limit / 10). Returns the status; x (n), y (m) are the scaled
// iterates the solver returns, iters the iteration count.
int ref_pdhg(int32_t m, int32_t n, const int64_t* start, const int32_t* index, const double* value, const double* c,
             const double* l, const double* u, const double* rl, const double* ru, int32_t simplex_iter_limit,
             int32_t threads, double* x, double* y, int64_t* iters, double* objective) {
    sov::Model M = make_model(m, n, start, index, value, c, l, u, rl, ru);
    sov::Options opt;
    opt.log_level = 0;
    opt.threads = threads;
    opt.simplex_iter_limit = simplex_iter_limit;
    sov::Deadline dl;
    auto r = sov::PdhgSolver::solve(M, opt, dl);
    std::memcpy(x, r.x.data(), sizeof(double) * r.x.size());
    std::memcpy(y, r.y.data(), sizeof(double) * r.y.size());
    *iters = r.iterations;
    *objective = r.objective;
    return static_cast<int>(r.status);
}

// Same, bounded by wall-clock time (the reference checks it every 256
// iterations); used to time the reference on a bounded sample.
int ref_pdhg_timed(int32_t m, int32_t n, const int64_t* start, const int32_t* index, const double* value,
                   const double* c, const double* l, const double* u, const double* rl, const double* ru,
                   double time_limit, int32_t threads, int64_t* iters, double* seconds) {
    sov::Model M = make_model(m, n, start, index, value, c, l, u, rl, ru);
    sov::Options opt;
    opt.log_level = 0;
    opt.threads = threads;
    sov::Deadline dl;
    dl.wall_limit_sec = time_limit;
    sov::Timer t;
    auto r = sov::PdhgSolver::solve(M, opt, dl);
    *seconds = t.elapsed_sec();
    *iters = r.iterations;
    return static_cast<int>(r.status);
}

int ref_scale(int32_t m, int32_t n, const int64_t* start, const int32_t* index, const double* value, const double* c,
              const double* l, const double* u, const double* rl, const double* ru, double* R, double* C) {
    sov::Model M = make_model(m, n, start, index, value, c, l, u, rl, ru);
    sov::ScalingInfo info;
    sov::scale_model(M, true, info);
    std::memcpy(R, info.R.data(), sizeof(double) * info.R.size());
    std::memcpy(C, info.C.data(), sizeof(double) * info.C.size());
    return info.active ? 1 : 0;
}

void ref_spmv(int32_t m, int32_t n, const int64_t* start, const int32_t* index, const double* value,
              const double* x, double* y) {
    sov::CscMatrix A(m, n);
    A.start.assign(start, start + n + 1);
    A.index.assign(index, index + start[n]);
    A.value.assign(value, value + start[n]);
    std::vector<double> xv(x, x + n), yv;
    A.spmv(xv, yv);
    std::memcpy(y, yv.data(), sizeof(double) * yv.size());
}

void ref_spmv_t(int32_t m, int32_t n, const int64_t* start, const int32_t* index, const double* value,
                const double* y, double* g) {
    sov::CscMatrix A(m, n);
    A.start.assign(start, start + n + 1);
    A.index.assign(index, index + start[n]);
    A.value.assign(value, value + start[n]);
    std::vector<double> yv(y, y + m), gv;
    A.spmv_transpose(yv, gv);
    std::memcpy(g, gv.data(), sizeof(double) * gv.size());
}

// Checker of the reference on x for the LP; returns 1 on pass.
int ref_verify(int32_t m, int32_t n, const int64_t* start, const int32_t* index, const double* value,
               const double* c, const double* l, const double* u, const double* rl, const double* ru,
               const double* x, double claimed, double tol, double* out3) {
    sov::Model M = make_model(m, n, start, index, value, c, l, u, rl, ru);
    std::vector<double> xv(x, x + n);
    auto rep = sov::verify_solution(M, xv, claimed, tol);
    out3[0] = rep.max_row_viol;
    out3[1] = rep.max_bound_viol;
    out3[2] = rep.computed_obj;
    return rep.passed ? 1 : 0;
}

// Presolver::presolve (src/presolve/presolve.hpp:32-345) at `level` on the LP
// with the flagged columns integer. ints = {status, solved, reduced nrow,
// reduced ncol, rows removed, cols removed, tightenings}; offset = reduced
// objective offset; the arrays as nir_lp_presolve returns them.
void ref_presolve(int32_t m, int32_t n, const int64_t* start, const int32_t* index, const double* value,
                  const double* c, const double* l, const double* u, const double* rl, const double* ru,
                  const int8_t* integer, int32_t level, int32_t* ints, double* offset, double* fixed_x,
                  int32_t* col_map, int32_t* row_map, double* collo, double* colup, double* rowlo, double* rowup) {
    sov::Model M = make_model(m, n, start, index, value, c, l, u, rl, ru);
    if (integer)
        for (int32_t j = 0; j < n; ++j)
            if (integer[j]) M.vartype[j] = sov::VarType::Integer;
    const sov::PresolveResult r = sov::Presolver::presolve(M, level);
    ints[0] = static_cast<int32_t>(r.status);
    ints[1] = r.solved_in_presolve ? 1 : 0;
    ints[2] = r.reduced.nrow;
    ints[3] = r.reduced.ncol;
    ints[4] = r.rows_removed;
    ints[5] = r.cols_removed;
    ints[6] = r.bound_tightenings;
    *offset = r.reduced.obj_offset;
    std::memcpy(fixed_x, r.fixed_x.data(), sizeof(double) * r.fixed_x.size());
    for (size_t j = 0; j < r.orig_col_map.size(); ++j) col_map[j] = r.orig_col_map[j];
    for (size_t i = 0; i < r.orig_row_map.size(); ++i) row_map[i] = r.orig_row_map[i];
    std::memcpy(collo, r.reduced.collo.data(), sizeof(double) * r.reduced.collo.size());
    std::memcpy(colup, r.reduced.colup.data(), sizeof(double) * r.reduced.colup.size());
    std::memcpy(rowlo, r.reduced.rowlo.data(), sizeof(double) * r.reduced.rowlo.size());
    std::memcpy(rowup, r.reduced.rowup.data(), sizeof(double) * r.reduced.rowup.size());
}

// The reference's simplex (LpSolver, src/lp/lp_solver.hpp) on the LP, from
// its slack basis (status == NULL) or from the given BasisSnapshot
// (status: n + m codes, basic: m variable numbers) through set_basis.
// Returns the status; iterations and objective through the pointers, x (n).
int ref_simplex(int32_t m, int32_t n, const int64_t* start, const int32_t* index, const double* value,
                const double* c, const double* l, const double* u, const double* rl, const double* ru,
                const int8_t* status, const int32_t* basic, int64_t* iters, double* objective, double* x) {
    sov::Model M = make_model(m, n, start, index, value, c, l, u, rl, ru);
    sov::Options opt;
    opt.log_level = 0;
    sov::LpSolver s(opt);
    s.load(M);
    if (status) {
        sov::BasisSnapshot snap;
        snap.status.resize(static_cast<size_t>(n) + m);
        for (int32_t v = 0; v < n + m; ++v) snap.status[v] = static_cast<sov::VarStatus>(status[v]);
        snap.basic_index.assign(basic, basic + m);
        s.set_basis(snap);
    }
    sov::Deadline dl;
    const sov::Status st = s.solve(dl);
    *iters = s.stats.iterations;
    *objective = s.objective;
    for (int32_t j = 0; j < n; ++j) x[j] = s.value[j];
    return static_cast<int>(st);
}

}  // extern "C"
