// Device evaluation of a point's relative KKT error, used by engine.hip for
// reference-mode sessions (the plain PDHG keeps no KKT terms of its own).
#pragma once
#include <hip/hip_runtime.h>

#include "devmat.hpp"

namespace nir {
namespace gpu {

// Relative KKT of (x, y) on min c^T x + offset, rl <= A x <= ru, lo <= x <= up
// (bounds as given; |b| >= 1e20 or non-finite = absent), with A given by its
// CSR copy (A) and the CSR copy of its transpose (AT), both in the segmented
// layout. out[0..4] = relative primal, relative dual, relative gap, primal
// objective, dual objective. Products A x and A^T y and every sum are taken in
// double-double. Synchronises `s`.
void kkt_device(hipStream_t s, const CsrShape& A, const double* aval, const CsrShape& AT, const double* atval,
                const double* x, const double* y, const double* c, const double* lo, const double* up,
                const double* rl, const double* ru, double offset, double* out);

}  // namespace gpu
}  // namespace nir
