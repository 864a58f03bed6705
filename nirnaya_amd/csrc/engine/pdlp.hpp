// Restarted PDHG (PDLP-style) engine interface, used by engine.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include <vector>

#include "../../../include/nirnaya/pdhg.h"
#include "pdlp_core.hpp"
#include "spmv.hpp"

namespace nir {
namespace gpu {

// Host double-double accumulator (two-sum / FMA two-product).
struct DDHost {
    double hi = 0.0, lo = 0.0;
    explicit DDHost(double h = 0.0) : hi(h) {}
    void add(double bh, double bl) {
        double s1 = hi + bh, v = s1 - hi, s2 = (hi - (s1 - v)) + (bh - v);
        double t1 = lo + bl, w = t1 - lo, t2 = (lo - (t1 - w)) + (bl - w);
        s2 += t1;
        double u = s1 + s2;
        s2 = s2 - (u - s1);
        s1 = u;
        s2 += t2;
        hi = s1 + s2;
        lo = s2 - (hi - s1);
    }
    void add_prod(double a, double b) {
        const double p = a * b;
        add(p, std::fma(a, b, -p));
    }
    double value() const { return hi + lo; }
};

struct Pdlp;
struct PdlpDeleter {
    void operator()(Pdlp* p) const;
};
Pdlp* pdlp_create();
struct DevMatrix;
void pdlp_setup(Pdlp& P, hipStream_t s, const nir_lp* lp, const nir_params& prm, DevMatrix* A, DevMatrix* AT);
void pdlp_run(Pdlp& P, hipStream_t s, const nir_params& prm, long long budget);
int pdlp_result(Pdlp& P, hipStream_t s, double* x, double* y, nir_info* info);
int pdlp_time(Pdlp& P, hipStream_t s, long long iters, nir_timing* out);
void pdlp_scaling(Pdlp& P, hipStream_t s, double* R, double* C);

// The matrix half of pdlp mode's set-up, shared by the single-LP engine and
// the family engine: scales the device matrices A and A^T in place (10 Ruiz
// passes, one Pock-Chambolle pass; factors accumulated into the device
// arrays dR, dC and returned in R, C) and returns the step scale
// eta = 0.998 / sigma_max of the scaled matrix (power iteration).
double pdlp_prepare_matrix(hipStream_t s, int m, int n, long long nnz, DevMatrix* A, DevMatrix* AT, double* dR,
                           double* dC, std::vector<double>& R, std::vector<double>& C);

// The instance half, on the host: an LP's costs and bounds in the scaled
// space, its caller-unit scales, its start point and its initial state.
struct PdlpInstance {
    std::vector<double> c, lo, up, rl, ru, rt, ct, x0, y0;
    PdlpState st;
};
void pdlp_prepare_instance(const nir_lp* lp, const nir_params& prm, const std::vector<double>& R,
                           const std::vector<double>& C, double eta, bool mixed, PdlpInstance& out);

// A run continued on this engine from a state made elsewhere (a lane of the
// family engine, between evaluation blocks): the scaled matrices, scaling and
// column bounds are shared, the member's vectors and state are copied on the
// device; the next pdlp_run starts with the next block. With `mixed` the
// state is the FP32 phase's (FP64 anchor in xa / ya, its products HA, GA,
// FP32 deltas and re-based constants); a state whose FP32 phase has ended
// switches to the FP64 tail here, as pdlp_run does.
struct PdlpAdopt {
    int m = 0, n = 0;
    long long nnz = 0;
    DevMatrix* A = nullptr;
    DevMatrix* AT = nullptr;
    const double *R = nullptr, *C = nullptr, *lo = nullptr, *up = nullptr;   // shared, device
    const double *c = nullptr, *rl = nullptr, *ru = nullptr;                  // the member's, device
    const double *x = nullptr, *xbar = nullptr, *xa = nullptr, *xp = nullptr;
    const double *y = nullptr, *ya = nullptr, *yp = nullptr;
    const PdlpState* st = nullptr;   // device
    std::vector<double> h_c;         // caller-unit costs (objective of the result)
    double offset = 0.0;
    bool mixed = false;
    const double *HA = nullptr, *GA = nullptr;
    const float *fx = nullptr, *fxbar = nullptr, *fxp = nullptr, *rc = nullptr, *dl = nullptr, *du = nullptr;
    const float *fy = nullptr, *fyp = nullptr, *drl = nullptr, *dru = nullptr, *sa = nullptr;
};
void pdlp_adopt(Pdlp& P, hipStream_t s, const nir_params& prm, const PdlpAdopt& a);

}  // namespace gpu
}  // namespace nir
