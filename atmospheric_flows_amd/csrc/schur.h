// Cellwise-direct pieces of the velocity block A (A is block diagonal:
// one (k+1)^2 x (k+1)^2 rho-weighted mass block per cell and component,
// euler_operator.h:2328-2357), used by the mixed form of the Schur
// complement and of the A solves (schur.hip).
#pragma once
#include "ops.h"

namespace atmo {

// y = A^{-1} (B p): B (vmult_pressure, euler_operator.h:2420-2435) with the
// cell blocks' FP64 inverses Ainv (symmetric, packed per cell: SymBlock) applied in
// the same pass; t = B p never leaves the chip.
template <int K>
void launch_b_ainv(const MeshDev& m, const OpConsts& oc, double coeff, const double* p, const double* Ainv,
                   double* y, long Ns, hipStream_t s);

// x = A^{-1} b cell by cell (both velocity components).
template <int K>
void launch_ainv_apply(const MeshDev& m, const double* Ainv, const double* b, double* x, long Ns, hipStream_t s);

}  // namespace atmo
