#include "core.hpp"

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <stdexcept>

namespace nir {

const char* status_name(Status s) {
    switch (s) {
        case Status::Optimal: return "optimal";
        case Status::Infeasible: return "infeasible";
        case Status::Unbounded: return "unbounded";
        case Status::InfeasibleOrUnbounded: return "infeasible_or_unbounded";
        case Status::TimeLimit: return "time_limit";
        case Status::WorkLimit: return "work_limit";
        case Status::IterationLimit: return "iteration_limit";
        case Status::NodeLimit: return "node_limit";
        case Status::SolutionLimit: return "solution_limit";
        case Status::GapLimit: return "gap_limit";
        case Status::Interrupted: return "interrupted";
        case Status::NumericalTrouble: return "numerical_trouble";
        case Status::InputError: return "input_error";
        case Status::NotSupported: return "not_supported";
        case Status::InternalError: return "internal_error";
        case Status::CutoffReached: return "cutoff_reached";
    }
    return "unknown";
}

// Process exit codes of the CLI (reference status.hpp:48-70).
int status_exit_code(Status s) {
    switch (s) {
        case Status::Optimal:
        case Status::GapLimit: return 0;
        case Status::TimeLimit:
        case Status::WorkLimit:
        case Status::IterationLimit:
        case Status::NodeLimit:
        case Status::SolutionLimit:
        case Status::Interrupted: return 1;
        case Status::Infeasible: return 2;
        case Status::Unbounded:
        case Status::InfeasibleOrUnbounded: return 3;
        case Status::NumericalTrouble: return 5;
        default: return 4;
    }
}

void Options::apply_preset(const std::string& p) {
    preset = p;
    if (p == "cautious" || p == "strict") {
        tol.primal_feas_tol = 1e-8;
        tol.dual_feas_tol = 1e-8;
        tol.lu_threshold_u = 0.5;
        refactor_interval = 30;
    } else if (p == "relaxed") {
        tol.primal_feas_tol = 1e-6;
        tol.dual_feas_tol = 1e-6;
    } else if (p == "fast") {
        presolve_rounds = 5;
    }
}

namespace {
// Whole-string numeric parses; trailing garbage is an error.
double parse_f64(const std::string& v) {
    size_t pos = 0;
    double d = std::stod(v, &pos);
    if (pos != v.size()) throw std::invalid_argument("trailing characters");
    return d;
}
long long parse_i64(const std::string& v) {
    size_t pos = 0;
    long long d = std::stoll(v, &pos);
    if (pos != v.size()) throw std::invalid_argument("trailing characters");
    return d;
}
int parse_i32(const std::string& v) {
    long long d = parse_i64(v);
    if (d < INT32_MIN || d > INT32_MAX) throw std::out_of_range("int32");
    return static_cast<int>(d);
}
bool truthy_on(const std::string& v) { return v == "1" || v == "true"; }
bool not_off(const std::string& v) { return v != "0" && v != "false"; }
}  // namespace

bool Options::set_by_name(const std::string& key, const std::string& val) {
    try {
        // Keys and value rules of the reference (options.hpp:61-89).
        if (key == "time_limit") { time_limit = parse_f64(val); return true; }
        if (key == "work_limit") { work_limit = parse_f64(val); return true; }
        if (key == "threads") { threads = parse_i32(val); return true; }
        if (key == "seed") { seed = static_cast<uint64_t>(std::stoull(val)); return true; }
        if (key == "log_level" || key == "log") { log_level = parse_i32(val); return true; }
        if (key == "deterministic") { deterministic = truthy_on(val); return true; }
        if (key == "method") { method = val; return true; }
        if (key == "preset") { apply_preset(val); return true; }
        if (key == "presolve") { presolve = parse_i32(val); return true; }
        if (key == "scale") { scale = not_off(val); return true; }
        if (key == "refactor_interval") { refactor_interval = parse_i32(val); return true; }
        if (key == "bound_flipping") { bound_flipping = not_off(val); return true; }
        if (key == "perturb") { perturb = not_off(val); return true; }
        if (key == "simplex_iter_limit") { simplex_iter_limit = parse_i32(val); return true; }
        if (key == "ipm_max_iter") { ipm_max_iter = parse_i32(val); return true; }
        if (key == "mip_gap") { tol.mip_gap = parse_f64(val); return true; }
        if (key == "mip_abs_gap") { tol.mip_abs_gap = parse_f64(val); return true; }
        if (key == "primal_feas_tol") { tol.primal_feas_tol = parse_f64(val); return true; }
        if (key == "dual_feas_tol") { tol.dual_feas_tol = parse_f64(val); return true; }
        if (key == "int_tol") { tol.int_tol = parse_f64(val); return true; }
        if (key == "lu_threshold") { tol.lu_threshold_u = parse_f64(val); return true; }
        if (key == "node_limit") { node_limit = parse_i32(val); return true; }
        if (key == "solution_limit") { solution_limit = parse_i32(val); return true; }
        if (key == "branching") { branching = val; return true; }
        if (key == "cuts") { enable_cuts = not_off(val); return true; }
        if (key == "heuristics") { enable_heuristics = not_off(val); return true; }
        // Keys added by this engine.
        if (key == "checker_tol") { tol.checker_tol = parse_f64(val); return true; }
        if (key == "presolve_rounds") { presolve_rounds = parse_i32(val); return true; }
        if (key == "sol" || key == "sol_path") { sol_path = val; return true; }
        if (key == "stats_json") { stats_json = val; return true; }
        if (key == "pdhg_mode") {
            if (val != "reference" && val != "pdlp") return false;
            pdhg.mode = val;
            return true;
        }
        if (key == "pdhg_precision") {
            if (val != "fp64" && val != "mixed") return false;
            pdhg.precision = val;
            return true;
        }
        if (key == "pdhg_tol") { pdhg.tol = parse_f64(val); return true; }
        if (key == "pdhg_polish_tol") { pdhg.polish_tol = parse_f64(val); return true; }
        if (key == "pdhg_host_scaling") { pdhg.host_scaling = parse_i32(val); return true; }
        if (key == "pdhg_max_iters") { pdhg.max_iters = parse_i64(val); return true; }
        if (key == "pdhg_eval_freq") {
            int f = parse_i32(val);
            if (f < 1) return false;
            pdhg.eval_freq = f;
            return true;
        }
        if (key == "pdhg_device") { pdhg.device = parse_i32(val); return true; }
        if (key == "pdhg_graph") { pdhg.use_graph = not_off(val) ? 1 : 0; return true; }
        if (key == "pdhg_verbose") { pdhg.verbose = parse_i32(val); return true; }
    } catch (const std::exception&) {
        return false;
    }
    return false;
}

void log_msg(int level_now, int min_level, const char* fmt, ...) {
    static std::mutex mu;
    if (level_now < min_level) return;
    std::lock_guard<std::mutex> lock(mu);
    va_list args;
    va_start(args, fmt);
    std::vprintf(fmt, args);
    va_end(args);
    std::fflush(stdout);
}

}  // namespace nir
