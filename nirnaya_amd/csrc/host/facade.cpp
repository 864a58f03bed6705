#include "facade.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>

#include "../../../include/nirnaya/pdhg.h"
#include "../engine/trace.hpp"

namespace nir {
namespace {

int deadline_poll(void* user) { return static_cast<const Deadline*>(user)->expired() ? 1 : 0; }

struct EngineRun {
    Status status = Status::InternalError;
    std::vector<double> x, y;
    long long iterations = 0;
    nir_info info{};
    std::string error;
};

// Hands the presolved, scaled, internal-min LP to the GPU engine; this is the
// call site of the reference's PdhgSolver::solve (solver.hpp:235-244).
EngineRun run_engine(const Model& lp, const Options& opt, const Deadline& deadline, const Scaling& sc,
                     const std::vector<double>* wx, const std::vector<double>* wy, double primal_weight = 0.0) {
    const TraceRange trace("sov engine");
    EngineRun r;
    nir_lp v{};
    v.nrow = lp.nrow;
    v.ncol = lp.ncol;
    v.col_start = lp.A.start.data();
    v.row_index = lp.A.index.data();
    v.value = lp.A.value.data();
    v.cost = lp.cost.data();
    v.collo = lp.collo.data();
    v.colup = lp.colup.data();
    v.rowlo = lp.rowlo.data();
    v.rowup = lp.rowup.data();
    v.obj_offset = lp.obj_offset;

    nir_params p;
    nir_params_default(&p);
    const bool pdlp = opt.pdhg.mode == "pdlp";
    p.mode = pdlp ? NIR_MODE_PDLP : NIR_MODE_REFERENCE;
    p.precision = (pdlp && opt.pdhg.precision == "mixed") ? NIR_PREC_MIXED : NIR_PREC_FP64;
    // Reference cap: max(5000, simplex_iter_limit / 10) (pdhg_solver.hpp:60).
    long long cap = std::max<long long>(5000, opt.simplex_iter_limit / 10);
    if (opt.pdhg.max_iters > 0) cap = opt.pdhg.max_iters;
    // The reference loop's counter is an Idx (int32, src/core/types.hpp:12).
    if (!pdlp) cap = std::min<long long>(cap, 2147483647LL);
    p.max_iters = cap;
    p.tol = opt.pdhg.tol;
    p.polish_tol = opt.pdhg.polish_tol;
    p.eval_freq = opt.pdhg.eval_freq;
    p.check_termination = 1;
    p.use_graph = opt.pdhg.use_graph;
    p.verbose = opt.pdhg.verbose;
    p.primal_weight = primal_weight;
    p.x0 = (wx && !wx->empty()) ? wx->data() : nullptr;
    p.y0 = (wy && !wy->empty()) ? wy->data() : nullptr;
    if (sc.active) {
        p.unscale_row = sc.row.data();
        p.unscale_col = sc.col.data();
    }
    p.should_stop = &deadline_poll;
    p.stop_user = const_cast<Deadline*>(&deadline);

    nir_ctx* ctx = nullptr;
    if (nir_ctx_create(opt.pdhg.device, &ctx) != 0) {
        r.error = std::string("GPU engine unavailable: ") + nir_last_error();
        return r;
    }
    r.x.assign(static_cast<size_t>(lp.ncol), 0.0);
    r.y.assign(static_cast<size_t>(lp.nrow), 0.0);
    const int rc = nir_pdhg_solve(ctx, &v, &p, r.x.data(), r.y.data(), &r.info);
    if (rc != 0) {
        r.error = std::string("GPU engine failed: ") + nir_ctx_error(ctx);
        nir_ctx_destroy(ctx);
        return r;
    }
    nir_ctx_destroy(ctx);
    r.status = static_cast<Status>(r.info.status);
    r.iterations = r.info.iterations;
    return r;
}

// Primal feasibility polish (after PDLP's feasibility polishing): the same
// engine on the feasibility problem - the LP with a zero objective - started
// from x with zero duals, in FP64 to a tight relative KKT. Starting from a
// point that already meets the optimality tolerance it reaches a nearby point
// whose every row and bound holds to ~1e-13 within a few thousand iterations
// and moves the objective by about the tolerance.
EngineRun polish_feasibility(const Model& lp, const Options& opt, const Deadline& deadline, const Scaling& sc,
                             const std::vector<double>& x) {
    Model feas = lp;
    std::fill(feas.cost.begin(), feas.cost.end(), 0.0);
    feas.obj_offset = 0.0;
    Options po = opt;
    po.pdhg.precision = "fp64";
    po.pdhg.polish_tol = 0.0;
    po.pdhg.tol = 1e-9;
    po.pdhg.max_iters = 20000;
    const std::vector<double> y0(static_cast<size_t>(lp.nrow), 0.0);
    return run_engine(feas, po, deadline, sc, &x, &y0);
}

// Large continuous LPs are certified on the device (nir_verify_solution: the
// same double-double activities and rules as verify_solution, names filled in
// here); smaller or integer models on the host.
constexpr long long kDeviceVerifyNnz = 1ll << 20;

CheckReport certify(const Model& m, const std::vector<double>& x, double claimed, double tol, int device) {
    const TraceRange trace("sov certify");
    if (m.A.nnz() < kDeviceVerifyNnz || m.num_integers() > 0 || static_cast<Idx>(x.size()) != m.ncol)
        return verify_solution(m, x, claimed, tol);
    CheckReport rep;
    nir_lp v{};
    v.nrow = m.nrow;
    v.ncol = m.ncol;
    v.col_start = m.A.start.data();
    v.row_index = m.A.index.data();
    v.value = m.A.value.data();
    v.cost = m.cost.data();
    v.collo = m.collo.data();
    v.colup = m.colup.data();
    v.rowlo = m.rowlo.data();
    v.rowup = m.rowup.data();
    v.obj_offset = m.obj_offset;
    nir_ctx* ctx = nullptr;
    nir_verify_report r{};
    if (nir_ctx_create(device, &ctx) != 0) {
        rep.passed = false;
        rep.worst_item = std::string("device certifier unavailable: ") + nir_last_error();
        return rep;
    }
    const int rc = nir_verify_solution(ctx, &v, x.data(), claimed, tol, &r);
    if (rc != 0) {
        rep.passed = false;
        rep.worst_item = std::string("device certifier failed: ") + nir_ctx_error(ctx);
        nir_ctx_destroy(ctx);
        return rep;
    }
    nir_ctx_destroy(ctx);
    if (r.nonfinite_col >= 0) {
        rep.passed = false;
        rep.worst_item = "NaN/Inf in variable " + m.col_names[r.nonfinite_col];
        return rep;
    }
    rep.max_bound_viol = r.max_bound_viol;
    rep.max_row_viol = r.max_row_viol;
    rep.computed_obj = r.computed_obj;
    rep.obj_rel_err = r.obj_rel_err;
    rep.passed = r.passed != 0;
    // The host checker names the worst violation above tol, rows after columns.
    if (r.max_row_viol > tol && r.worst_row >= 0)
        rep.worst_item = (r.worst_row_upper ? "Row upper bound on " : "Row lower bound on ") + m.row_names[r.worst_row];
    else if (r.max_bound_viol > tol && r.worst_col >= 0)
        rep.worst_item = (r.worst_col_upper ? "Upper bound on " : "Lower bound on ") + m.col_names[r.worst_col];
    return rep;
}

}  // namespace

SolveResult Solver::solve(const Model& user_model) {
    SolveResult out;
    stop_flag.store(false, std::memory_order_relaxed);
    const int L = options.log_level;
    Deadline deadline;
    deadline.wall_limit_sec = options.time_limit;
    deadline.work_limit = options.work_limit;
    deadline.stop_flag = &stop_flag;

    Model valid = user_model;
    const Status vs = valid.validate(out.message);
    if (vs != Status::Optimal) {
        out.status = vs;
        out.solve_time_sec = deadline.elapsed();
        return out;
    }
    log_msg(L, 1, "========================================================================\n");
    log_msg(L, 1, "  NIRNAYA-MI355X | model %s | rows %d | cols %d | nnz %lld | ints %d | qnnz %lld\n",
            valid.name.c_str(), valid.nrow, valid.ncol, static_cast<long long>(valid.A.nnz()), valid.num_integers(),
            static_cast<long long>(valid.Q.nnz()));
    log_msg(L, 1, "========================================================================\n");

    const Model internal = valid.to_internal_min();
    Presolved pre = [&] {
        const TraceRange trace("sov presolve");
        return presolve(internal, options.presolve, options.presolve_rounds);
    }();
    const double t_presolve = deadline.elapsed();
    if (options.presolve > 0 && (pre.rows_removed > 0 || pre.cols_removed > 0 || pre.tightenings > 0))
        log_msg(L, 1, "Presolve: %d rows, %d cols remaining (removed %d rows, %d cols, %d tightenings)\n",
                pre.reduced.nrow, pre.reduced.ncol, pre.rows_removed, pre.cols_removed, pre.tightenings);
    if (pre.solved) {
        out.status = pre.status;
        if (pre.status == Status::Optimal) {
            pre.postsolve({}, out.x);
            out.has_incumbent = true;
            out.objective = valid.eval_objective(out.x);
            out.best_bound = out.objective;
            out.verification = verify_solution(valid, out.x, out.objective, options.tol.checker_tol);
        }
        out.solve_time_sec = deadline.elapsed();
        return out;
    }

    // The reference's geometric pre-scaling (scaling.hpp) precedes its plain
    // loop; the pdlp mode scales inside the engine (Ruiz + Pock-Chambolle),
    // and stacking both slowed convergence on most LPs measured (the
    // 2,000 x 5,000 equality LP 6.7x more iterations, the 400k x 1.2M one
    // > 3.8x), so there it only runs when pdhg_host_scaling asks for it.
    const bool host_scale = options.scale && (options.pdhg.mode != "pdlp" || options.pdhg.host_scaling != 0);
    Scaling sc;
    const Model scaled = [&] {
        const TraceRange trace("sov scaling");
        return scale_geometric(pre.reduced, host_scale, sc);
    }();
    const double t_scaled = deadline.elapsed();
    double t_engine = t_scaled;
    std::vector<double> x_red(static_cast<size_t>(pre.reduced.ncol), 0.0);
    std::vector<double> y_red(static_cast<size_t>(pre.reduced.nrow), 0.0);

    if (scaled.is_qp() || scaled.is_mip()) {
        out.status = Status::NotSupported;
        out.message = scaled.is_qp() ? "Quadratic objectives are not handled by the PDHG engine"
                                     : "Integer variables are not handled by the PDHG engine";
    } else if (options.method != "pdhg" && options.method != "gpu") {
        out.status = Status::NotSupported;
        out.message = "method '" + options.method + "' is not built; use method=pdhg";
    } else {
        log_msg(L, 1, "Solving LP via GPU PDHG (%s, %s) on MI355X...\n", options.pdhg.mode.c_str(),
                options.pdhg.precision.c_str());
        // Warm start in the reduced, scaled space.
        std::vector<double> wx, wy;
        if (start_x.size() == static_cast<size_t>(valid.ncol)) {
            wx.resize(static_cast<size_t>(scaled.ncol));
            for (Idx k = 0; k < scaled.ncol; ++k) wx[k] = start_x[pre.col_map[k]] / (sc.active ? sc.col[k] : 1.0);
        }
        if (start_y.size() == static_cast<size_t>(valid.nrow)) {
            const double sgn = valid.sense == Sense::Maximize ? -1.0 : 1.0;
            wy.resize(static_cast<size_t>(scaled.nrow));
            for (Idx k = 0; k < scaled.nrow; ++k) wy[k] = sgn * start_y[pre.row_map[k]] / (sc.active ? sc.row[k] : 1.0);
        }
        Options eo = options;
        EngineRun er = run_engine(scaled, eo, deadline, sc, &wx, &wy);
        long long iterations = 0;
        for (int tightened = 0;; ++tightened) {
            if (!er.error.empty()) {
                out.status = Status::InternalError;
                out.message = er.error;
                break;
            }
            iterations += er.iterations;
            out.status = er.status;
            out.iterations = static_cast<Idx>(std::min<long long>(iterations, INT32_MAX));
            out.engine_iterations += er.iterations + er.info.polish_iterations;
            out.engine_seconds += er.info.setup_seconds + er.info.solve_seconds;
            out.rel_primal = er.info.rel_primal;
            out.rel_dual = er.info.rel_dual;
            out.rel_gap = er.info.rel_gap;
            out.restarts = er.info.restarts;
            out.primal_weight = er.info.primal_weight;
            if (out.status != Status::Optimal) break;
            sc.unscale_primal(er.x, x_red);
            sc.unscale_duals(er.y, y_red);
            out.has_incumbent = true;
            // pdlp mode: a relative KKT tolerance (l2 norms over all rows and
            // columns) does not bound each row's violation, which the
            // certificate below checks against checker_tol. When it would
            // fail: a primal feasibility polish from the engine's point; if
            // that fails too, continue from the point and primal weight with
            // the tolerance tightened tenfold (at most twice).
            if (eo.pdhg.mode != "pdlp") break;
            std::vector<double> xo;
            pre.postsolve(x_red, xo);
            const CheckReport chk = certify(valid, xo, valid.eval_objective(xo), options.tol.checker_tol,
                                            options.pdhg.device);
            if (chk.passed) break;
            // First a primal feasibility polish from this point; the duals
            // stay those of the optimality run.
            const EngineRun fp = polish_feasibility(scaled, eo, deadline, sc, er.x);
            if (fp.error.empty() && (fp.status == Status::Optimal || fp.status == Status::IterationLimit)) {
                std::vector<double> xr, xo2;
                sc.unscale_primal(fp.x, xr);
                pre.postsolve(xr, xo2);
                const double o1 = valid.eval_objective(xo), o2 = valid.eval_objective(xo2);
                const CheckReport c2 = certify(valid, xo2, o2, options.tol.checker_tol, options.pdhg.device);
                out.engine_iterations += fp.iterations;
                out.engine_seconds += fp.info.setup_seconds + fp.info.solve_seconds;
                log_msg(L, 1, "Feasibility polish: %lld iterations, max row violation %.2e -> %.2e, objective "
                              "moved by %.2e relative\n", static_cast<long long>(fp.iterations), chk.max_row_viol,
                        c2.max_row_viol, std::fabs(o2 - o1) / (1.0 + std::fabs(o1)));
                if (c2.passed) {
                    x_red = xr;
                    break;
                }
            }
            if (tightened == 2) break;
            eo.pdhg.tol /= 10.0;
            log_msg(L, 1, "Certificate would fail (max row violation %.2e, bound %.2e > %.0e): continuing to "
                          "relative KKT %.0e\n", chk.max_row_viol, chk.max_bound_viol, options.tol.checker_tol,
                    eo.pdhg.tol);
            const std::vector<double> cx = er.x, cy = er.y;
            er = run_engine(scaled, eo, deadline, sc, &cx, &cy, er.info.primal_weight);
        }
        t_engine = deadline.elapsed();
    }

    if (out.has_incumbent) {
        pre.postsolve(x_red, out.x);
        valid.A.multiply(out.x, out.row_activity);
        out.objective = valid.eval_objective(out.x);
        out.best_bound = out.objective;
        out.gap = 0.0;
        const double sgn = valid.sense == Sense::Maximize ? -1.0 : 1.0;
        out.row_duals.assign(static_cast<size_t>(valid.nrow), 0.0);
        for (size_t i = 0; i < pre.row_map.size() && i < y_red.size(); ++i) out.row_duals[pre.row_map[i]] = sgn * y_red[i];
        out.reduced_costs.assign(static_cast<size_t>(valid.ncol), 0.0);

        out.verification = certify(valid, out.x, out.objective, options.tol.checker_tol, options.pdhg.device);
        if (!out.verification.passed && out.status == Status::Optimal) {
            // The reference's fallback (a cold re-solve without presolve and
            // scaling); pdlp mode has already tightened its tolerance instead.
            if ((options.presolve > 0 || options.scale) && options.pdhg.mode != "pdlp") {
                Options fo = options;
                fo.presolve = 0;
                fo.scale = false;
                fo.log_level = 0;
                Solver fb_solver(fo);
                fb_solver.start_x = start_x;
                fb_solver.start_y = start_y;
                SolveResult fb = fb_solver.solve(user_model);
                if (fb.verification.passed && fb.status == Status::Optimal) {
                    fb.solve_time_sec = deadline.elapsed();
                    return fb;
                }
            }
            out.status = Status::NumericalTrouble;
            out.message = "Verification residual exceeded tolerance on " + out.verification.worst_item;
        }
    }
    out.solve_time_sec = deadline.elapsed();
    log_msg(L, 2, "  Timing:     validate+presolve %.3f s | scaling %.3f s | engine %.3f s | postsolve+certify %.3f s\n",
            t_presolve, t_scaled - t_presolve, t_engine - t_scaled, out.solve_time_sec - t_engine);

    log_msg(L, 1, "------------------------------------------------------------------------\n");
    log_msg(L, 1, "  Status:     %s\n", status_name(out.status));
    if (!out.message.empty()) log_msg(L, 1, "  Message:    %s\n", out.message.c_str());
    if (out.has_incumbent) {
        log_msg(L, 1, "  Objective:  %.12g\n", out.objective);
        log_msg(L, 1, "  Check (dd): PASS=%s | MaxRowViol=%.2e | MaxBndViol=%.2e | MaxIntViol=%.2e\n",
                out.verification.passed ? "YES" : "NO", out.verification.max_row_viol, out.verification.max_bound_viol,
                out.verification.max_int_viol);
    }
    log_msg(L, 1, "  Iterations: %d | Time: %.4f s\n", out.iterations, out.solve_time_sec);
    if (out.engine_iterations > 0 && options.pdhg.mode == "pdlp")
        log_msg(L, 1, "  PDHG:       rel. KKT primal %.2e dual %.2e gap %.2e | restarts %d | primal weight %.4g\n",
                out.rel_primal, out.rel_dual, out.rel_gap, out.restarts, out.primal_weight);
    if (out.engine_iterations > 0 && out.engine_seconds > 0.0)
        log_msg(L, 1, "  Engine:     %lld iterations in %.4f s (%.0f iterations/s, set-up included)\n",
                out.engine_iterations, out.engine_seconds, out.engine_iterations / out.engine_seconds);
    log_msg(L, 1, "========================================================================\n");

    if (!options.sol_path.empty() && out.has_incumbent)
        write_sol(options.sol_path, valid, out.status, out.objective, out.gap, out.x);
    if (!options.stats_json.empty()) write_stats_json(options.stats_json, valid, out);
    return out;
}

bool Solver::write_stats_json(const std::string& path, const Model& m, const SolveResult& r) {
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    auto fin = [](double v) { return std::isfinite(v) ? v : 0.0; };
    std::fprintf(f, "{\n");
    std::fprintf(f, "  \"solver\": \"NIRNAYA-MI355X\",\n");
    std::fprintf(f, "  \"model\": \"%s\",\n", m.name.c_str());
    std::fprintf(f, "  \"status\": \"%s\",\n", status_name(r.status));
    std::fprintf(f, "  \"objective\": %.17g,\n", fin(r.objective));
    std::fprintf(f, "  \"best_bound\": %.17g,\n", fin(r.best_bound));
    std::fprintf(f, "  \"gap\": %.17g,\n", fin(r.gap));
    std::fprintf(f, "  \"iterations\": %d,\n", r.iterations);
    std::fprintf(f, "  \"nodes\": %d,\n", r.nodes);
    std::fprintf(f, "  \"cuts_added\": %d,\n", r.cuts_added);
    std::fprintf(f, "  \"solve_time_sec\": %.9g,\n", r.solve_time_sec);
    std::fprintf(f, "  \"engine_iterations\": %lld,\n", r.engine_iterations);
    std::fprintf(f, "  \"engine_seconds\": %.9g,\n", r.engine_seconds);
    std::fprintf(f, "  \"rel_kkt\": [%.9g, %.9g, %.9g],\n", fin(r.rel_primal), fin(r.rel_dual), fin(r.rel_gap));
    std::fprintf(f, "  \"restarts\": %d,\n", r.restarts);
    std::fprintf(f, "  \"primal_weight\": %.9g,\n", fin(r.primal_weight));
    std::fprintf(f, "  \"verified\": %s\n", r.verification.passed ? "true" : "false");
    std::fprintf(f, "}\n");
    std::fclose(f);
    return true;
}

}  // namespace nir
