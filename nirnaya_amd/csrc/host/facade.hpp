// Solve pipeline around the GPU PDHG engine. Follows the order of the
// reference facade (include/sov/solver.hpp:128-326) for continuous LPs sent
// to method "pdhg" / "gpu":
//   validate -> to_internal_min -> presolve -> geometric scaling -> engine
//   -> unscale -> postsolve -> row activity, dd objective, row duals
//   -> dd verification -> (on failure) re-solve without presolve/scaling
//   -> NumericalTrouble.
// Models this engine does not solve (QP, MIP, other methods) return
// NotSupported with a message naming what was asked.
#pragma once
#include <atomic>
#include <string>
#include <vector>

#include "prep.hpp"

namespace nir {

struct SolveResult {
    Status status = Status::InternalError;
    double objective = kInf;
    double best_bound = -kInf;
    double gap = 0.0;
    bool has_incumbent = false;
    std::vector<double> x;
    std::vector<double> row_activity;
    std::vector<double> row_duals;
    std::vector<double> reduced_costs;
    Idx iterations = 0;
    Idx nodes = 0;
    Idx cuts_added = 0;
    double solve_time_sec = 0.0;
    CheckReport verification;
    std::string message;
    // Engine detail (not in the reference result).
    long long engine_iterations = 0;
    double engine_seconds = 0.0;
    double rel_primal = 0.0, rel_dual = 0.0, rel_gap = 0.0;
    int restarts = 0;             // pdlp mode: restarts of the last engine run
    double primal_weight = 0.0;   // pdlp mode: its final primal weight
};

class Solver {
public:
    Options options;
    std::atomic<bool> stop_flag{false};
    // Optional warm start for the PDHG engine, in the user's column / row space.
    std::vector<double> start_x, start_y;

    explicit Solver(const Options& o = Options()) : options(o) {}
    void interrupt() { stop_flag.store(true, std::memory_order_relaxed); }
    SolveResult solve(const Model& user_model);

    static bool write_stats_json(const std::string& path, const Model& m, const SolveResult& r);
};

}  // namespace nir
