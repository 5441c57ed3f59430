// Geometric multigrid pieces on the structured level hierarchy (mg.hip):
// DG embedding transfer (deal.II MGTransferMatrixFree for FE_DGQ, used at
// Atmospheric_Flow.cc:475-476 and its copies), L2 projection of the level
// densities (interpolate_to_mg, :1127-1132), Chebyshev smoother updates
// (PreconditionChebyshev, :489-503) and the coarse CG (MGCoarseGridIterativeSolver
// with identity preconditioner, :505-517).
#pragma once
#include "common.h"
#include "krylov.h"

namespace atmo {

// coarse = P^T fine (overwrites coarse)
template <int K, typename T>
void launch_restrict(const MeshDev& mf, const MeshDev& mc, const T* fine, T* coarse, int ncomp, long Nf, long Nc,
                     hipStream_t s);
// coarse = L2 projection of fine (density)
template <int K>
void launch_project_down(const MeshDev& mf, const MeshDev& mc, const double* fine, double* coarse, hipStream_t s);

// Coefficients of one Chebyshev sweep: step s is d = f1[s] d + f2[s] P (b - A x), x += d.
struct ChebSteps {
  static constexpr int MAX = 16;
  int n = 0;
  double f1[MAX] = {}, f2[MAX] = {};
};

// One Chebyshev sweep of a level from a zero guess (the V-cycle's
// pre-smoothing) fused with the residual it leaves: x = S(b), res = b - A x.
// The mass-type operators are cell-local, so every step runs in one kernel.
template <int K, typename T>
void launch_presmooth_residual(const MeshDev& m, int ncomp, bool weighted, T* x, const T* rho, T coeff, const T* b,
                               const T* P, const ChebSteps& cs, T* res, long Ns, hipStream_t s);
// x += prolongation of the coarse level's xc, then one Chebyshev sweep from
// that guess (the V-cycle's coarse-grid correction and post-smoothing).
template <int K, typename T>
void launch_prolongate_postsmooth(const MeshDev& m, const MeshDev& mc, int ncomp, bool weighted, T* x, const T* xc,
                                  const T* rho, T coeff, const T* b, const T* P, const ChebSteps& cs, long Ns,
                                  long Nc, hipStream_t s);

// Single-workgroup CG on the coarsest level (FP64, identity preconditioner,
// absolute tolerance read from *tol, at most max_its iterations). Work
// vectors g, d, h have the size of x. Writes the iteration count (or -1 on
// no convergence) to *its.
template <int K>
void launch_coarse_cg(const MeshDev& m, int ncomp, bool weighted, const double* rho, double coeff, const double* b,
                      double* x, double* g, double* d, double* h, const double* tol, int max_its, int* its, long Ns,
                      hipStream_t s);

}  // namespace atmo
