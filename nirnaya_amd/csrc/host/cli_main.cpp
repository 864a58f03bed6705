// sovsolve <model.mps|model.nlpb> [options] - command line front end of the facade.
// Flags and exit codes follow src/driver/cli_main.cpp:35-98 and
// src/core/status.hpp:48-70. Every "--flag value" is also accepted as
// "--flag=value" (so --method=pdhg works), and "--set key=value" reaches any
// option by name.
#include <cstdio>
#include <string>

#include "facade.hpp"

namespace {

void usage() {
    std::printf(
        "Usage: sovsolve <model.mps|model.nlpb> [options]\n"
        "  --method <m>              pdhg | gpu (GPU PDHG engine)\n"
        "  --pdhg-mode <m>           reference | pdlp\n"
        "  --precision <p>           fp64 | mixed (pdlp mode)\n"
        "  --tol <eps>               relative KKT tolerance (pdlp mode)\n"
        "  --sol <path>              write the solution (.sol)\n"
        "  --stats-json <path>       write solve statistics (.json)\n"
        "  --time-limit <sec>        wall-clock limit (default 600)\n"
        "  --threads <n>             accepted for compatibility\n"
        "  --presolve <0|1|2>        presolve level (default 2)\n"
        "  --seed <n>                random seed\n"
        "  --log <0|1|2|3>           log level (default 1)\n"
        "  --no-scale                disable geometric-mean scaling\n"
        "  --deterministic           accepted for compatibility\n"
        "  --tolerance-preset <p>    strict | default | relaxed\n"
        "  --set <name>=<value>      set any option by name\n"
        "  --write-lpb <path>        write the model as a binary LP (.nlpb) and exit\n"
        "  --write-mps <path>        write the model as MPS and exit\n"
        "  --presolve-report         run validation and presolve only, print what they removed\n"
        "  --version                 print the version\n");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        usage();
        return 4;
    }
    nir::Options opt;
    std::string path, lpb_out, mps_out;
    bool presolve_report = false;
    try {
        for (int i = 1; i < argc; ++i) {
            std::string arg = argv[i];
            std::string val;
            bool has_inline = false;
            if (arg.rfind("--", 0) == 0) {
                const size_t eq = arg.find('=');
                if (eq != std::string::npos && arg.rfind("--set", 0) != 0) {
                    val = arg.substr(eq + 1);
                    arg = arg.substr(0, eq);
                    has_inline = true;
                }
            }
            auto value = [&](std::string& out) {
                if (has_inline) {
                    out = val;
                    return true;
                }
                if (i + 1 < argc) {
                    out = argv[++i];
                    return true;
                }
                return false;
            };
            std::string v;
            if (arg == "--help" || arg == "-h") {
                usage();
                return 0;
            } else if (arg == "--version") {
                std::printf("sovsolve (NIRNAYA-MI355X PDHG engine) 1.0.0\n");
                return 0;
            } else if (arg == "--no-scale") {
                opt.scale = false;
            } else if (arg == "--deterministic") {
                opt.deterministic = true;
            } else if (arg == "--presolve-report") {
                presolve_report = true;
            } else if (arg == "--write-lpb" && value(v)) {
                lpb_out = v;
            } else if (arg == "--write-mps" && value(v)) {
                mps_out = v;
            } else if (arg == "--sol" && value(v)) {
                opt.sol_path = v;
            } else if (arg == "--stats-json" && value(v)) {
                opt.stats_json = v;
            } else if (arg == "--time-limit" && value(v)) {
                opt.set_by_name("time_limit", v);
            } else if (arg == "--threads" && value(v)) {
                opt.set_by_name("threads", v);
            } else if (arg == "--method" && value(v)) {
                opt.method = v;
            } else if (arg == "--pdhg-mode" && value(v)) {
                opt.set_by_name("pdhg_mode", v);
            } else if (arg == "--precision" && value(v)) {
                opt.set_by_name("pdhg_precision", v);
            } else if (arg == "--tol" && value(v)) {
                opt.set_by_name("pdhg_tol", v);
            } else if (arg == "--presolve" && value(v)) {
                opt.set_by_name("presolve", v);
            } else if (arg == "--mip-gap" && value(v)) {
                opt.set_by_name("mip_gap", v);
            } else if (arg == "--mip-abs-gap" && value(v)) {
                opt.set_by_name("mip_abs_gap", v);
            } else if (arg == "--node-limit" && value(v)) {
                opt.set_by_name("node_limit", v);
            } else if (arg == "--seed" && value(v)) {
                opt.set_by_name("seed", v);
            } else if (arg == "--log" && value(v)) {
                opt.set_by_name("log_level", v);
            } else if (arg == "--tolerance-preset" && value(v)) {
                opt.apply_preset(v);
            } else if (arg.rfind("--set", 0) == 0) {
                std::string kv;
                if (arg.size() > 6 && arg[5] == '=') kv = arg.substr(6);
                else if (i + 1 < argc) kv = argv[++i];
                const size_t eq = kv.find('=');
                if (eq != std::string::npos && !opt.set_by_name(kv.substr(0, eq), kv.substr(eq + 1)))
                    std::fprintf(stderr, "Warning: option %s not applied\n", kv.c_str());
            } else if (!arg.empty() && arg[0] != '-') {
                path = arg;
            }
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error parsing arguments: %s\n", e.what());
        return 4;
    }
    if (path.empty()) {
        std::fprintf(stderr, "Error: No input MPS model specified.\n");
        return 4;
    }
    nir::Model model;
    std::string err;
    const nir::Status rs = nir::read_model(path, model, err);
    if (rs != nir::Status::Optimal) {
        std::fprintf(stderr, "Input Error reading %s: %s\n", path.c_str(), err.c_str());
        return nir::status_exit_code(rs);
    }
    if (!lpb_out.empty() || !mps_out.empty()) {
        if (!lpb_out.empty() && !nir::write_lpb(lpb_out, model, err)) {
            std::fprintf(stderr, "Error writing %s: %s\n", lpb_out.c_str(), err.c_str());
            return 4;
        }
        if (!mps_out.empty() && !nir::write_mps(mps_out, model)) {
            std::fprintf(stderr, "Error writing %s\n", mps_out.c_str());
            return 4;
        }
        std::printf("Wrote %s%s%s (%d rows, %d columns, %lld nonzeros)\n", lpb_out.c_str(),
                    (!lpb_out.empty() && !mps_out.empty()) ? " and " : "", mps_out.c_str(), model.nrow, model.ncol,
                    static_cast<long long>(model.A.nnz()));
        return 0;
    }
    if (presolve_report) {
        // The facade's steps before the engine (solver.hpp:156-187 order):
        // validate, internal min form, presolve.
        std::string msg;
        const nir::Status vs = model.validate(msg);
        if (vs != nir::Status::Optimal) {
            std::printf("validation: %s (%s)\n", nir::status_name(vs), msg.c_str());
            return nir::status_exit_code(vs);
        }
        const nir::Presolved pre = nir::presolve(model.to_internal_min(), opt.presolve, opt.presolve_rounds);
        std::printf("presolve level %d: rows %d -> %d, cols %d -> %d, removed %d rows, %d cols, %d tightenings%s\n",
                    opt.presolve, model.nrow, pre.reduced.nrow, model.ncol, pre.reduced.ncol, pre.rows_removed,
                    pre.cols_removed, pre.tightenings, pre.solved ? ", solved" : "");
        return 0;
    }
    nir::Solver solver(opt);
    const nir::SolveResult res = solver.solve(model);
    return nir::status_exit_code(res.status);
}
