// Core vocabulary shared by the host shell and the engine boundary: index
// types, the infinity convention, solve status codes, the string-keyed option
// set, the wall-clock/interrupt deadline and double-double arithmetic.
//
// Semantics follow the reference's core headers:
//   src/core/types.hpp:12-16,49-61  (Idx/Off, 1e20 infinity, bound sanitising)
//   src/core/status.hpp:7-70        (status codes, names, CLI exit codes)
//   src/core/options.hpp:9-89       (defaults and set_by_name keys)
//   src/core/timer.hpp:41-47        (Deadline::expired)
//   src/core/dd_real.hpp:8-104      (two-sum / two-prod double-double)
#pragma once
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <limits>
#include <string>
#include <vector>

namespace nir {

using Idx = int32_t;   // row / column index
using Off = int64_t;   // offset into nonzero arrays

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kInfInput = 1e20;   // |b| >= 1e20 means "no bound"

enum class VarType : int8_t { Continuous = 0, Integer = 1, Binary = 2, SemiCont = 3 };
enum class Sense : int8_t { Minimize = 1, Maximize = -1 };

// A bound is real only when it is finite and below the 1e20 input infinity
// (types.hpp:49-51).
inline bool finite_bound(double b) { return std::isfinite(b) && std::fabs(b) < kInfInput; }
inline double clean_lower(double b) { return (!std::isfinite(b) || b <= -kInfInput) ? -kInf : b; }
inline double clean_upper(double b) { return (!std::isfinite(b) || b >= kInfInput) ? kInf : b; }

// std::max / std::min exactly as <algorithm> defines them, spelled out so the
// NaN and signed-zero behaviour is visible: max(a,b) = (a<b)?b:a,
// min(a,b) = (b<a)?b:a.
inline double amax(double a, double b) { return (a < b) ? b : a; }
inline double amin(double a, double b) { return (b < a) ? b : a; }

// ---------------------------------------------------------------- status
// Integer values are part of the C ABI (sov_get_status) and must not move.
enum class Status : int32_t {
    Optimal = 0, Infeasible = 1, Unbounded = 2, InfeasibleOrUnbounded = 3,
    TimeLimit = 4, WorkLimit = 5, IterationLimit = 6, NodeLimit = 7,
    SolutionLimit = 8, GapLimit = 9, Interrupted = 10, NumericalTrouble = 11,
    InputError = 12, NotSupported = 13, InternalError = 14, CutoffReached = 15
};

const char* status_name(Status s);
int status_exit_code(Status s);

// ---------------------------------------------------------------- options
struct Tolerances {
    double primal_feas_tol = 1e-7;
    double dual_feas_tol = 1e-7;
    double int_tol = 1e-6;
    double lu_threshold_u = 0.1;
    double mip_gap = 1e-4;
    double mip_abs_gap = 1e-6;
    double checker_tol = 1e-6;
};

// PDHG engine selection. "reference" reproduces the reference's fixed-step,
// unrestarted loop iteration for iteration; "pdlp" is the restarted engine.
struct PdhgOptions {
    std::string mode = "reference";   // reference | pdlp
    std::string precision = "fp64";   // fp64 | mixed
    double tol = 1e-6;                // relative KKT target (pdlp mode)
    double polish_tol = 0.0;          // >0: FP64 polish target after mixed phase
    long long max_iters = 0;          // 0: derive from simplex_iter_limit / defaults
    int eval_freq = 64;               // iterations between KKT evaluations (pdlp)
    int device = 0;
    int use_graph = 1;                // capture iteration blocks in hipGraphs
    int verbose = 0;
    // pdlp mode: also apply the reference's geometric pre-scaling before the
    // engine's own Ruiz + Pock-Chambolle scaling (0 = engine scaling only).
    int host_scaling = 0;
};

struct Options {
    double time_limit = 600.0;
    double work_limit = kInf;
    int threads = 1;
    uint64_t seed = 1;
    int log_level = 1;
    bool deterministic = true;
    int memory_limit_mb = 0;
    std::string method = "auto";
    std::string preset = "default";
    std::string stats_json;
    std::string sol_path;
    Tolerances tol;
    int presolve = 2;
    int presolve_rounds = 20;
    bool scale = true;
    int refactor_interval = 100;
    bool bound_flipping = true;
    bool perturb = true;
    Idx simplex_iter_limit = 500000;
    int ipm_max_iter = 200;
    std::string branching = "reliability";
    Idx node_limit = 100000;
    Idx solution_limit = 0;
    bool enable_cuts = true;
    bool enable_heuristics = true;
    PdhgOptions pdhg;

    void apply_preset(const std::string& p);
    // Returns false for an unknown key or a value that does not parse
    // (the reference lets std::stod exceptions escape; here they are caught).
    bool set_by_name(const std::string& key, const std::string& val);
};

// ---------------------------------------------------------------- deadline
struct Deadline {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double wall_limit_sec = kInf;
    double work_limit = kInf;
    int64_t work = 0;
    const std::atomic<bool>* stop_flag = nullptr;

    double elapsed() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    bool expired() const {
        if (stop_flag && stop_flag->load(std::memory_order_relaxed)) return true;
        if (elapsed() >= wall_limit_sec) return true;
        if (static_cast<double>(work) >= work_limit) return true;
        return false;
    }
};

// ---------------------------------------------------------------- double-double
// Error-free transformations (Knuth two-sum, Dekker quick-two-sum, FMA
// two-product). The operation order is fixed by the algorithms themselves.
struct DD {
    double hi = 0.0, lo = 0.0;
    DD() = default;
    explicit DD(double h, double l = 0.0) : hi(h), lo(l) {}

    static inline void two_sum(double a, double b, double& s, double& e) {
        s = a + b;
        double bv = s - a;
        e = (a - (s - bv)) + (b - bv);
    }
    static inline void fast_two_sum(double a, double b, double& s, double& e) {
        s = a + b;
        e = b - (s - a);
    }
    void add(double b) {
        double s, e;
        two_sum(hi, b, s, e);
        e += lo;
        fast_two_sum(s, e, hi, lo);
    }
    void add(const DD& b) {
        double s1, s2, t1, t2;
        two_sum(hi, b.hi, s1, s2);
        two_sum(lo, b.lo, t1, t2);
        s2 += t1;
        fast_two_sum(s1, s2, s1, s2);
        s2 += t2;
        fast_two_sum(s1, s2, hi, lo);
    }
    // this += a*b with the product kept exact (p + e == a*b).
    void add_prod(double a, double b) {
        double p = a * b;
        double e = std::fma(a, b, -p);
        add(DD(p, e));
    }
    double value() const { return hi + lo; }
};

// Neumaier compensated summation.
struct CompensatedSum {
    double s = 0.0, c = 0.0;
    void add(double x) {
        double t = s + x;
        if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    double value() const { return s + c; }
};

// printf-style logger honouring the log_level option.
void log_msg(int level_now, int min_level, const char* fmt, ...);

}  // namespace nir
