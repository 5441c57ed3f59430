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
// fine += P coarse
template <int K, typename T>
void launch_prolongate_add(const MeshDev& mf, const MeshDev& mc, const T* coarse, T* fine, int ncomp, long Nf,
                           long Nc, hipStream_t s);
// coarse = L2 projection of fine (density)
template <int K>
void launch_project_down(const MeshDev& mf, const MeshDev& mc, const double* fine, double* coarse, hipStream_t s);

// Chebyshev updates (P = inverse diagonal): the first step from a zero
// guess, d = P b / theta ; x = d (the later steps need A x and are fused
// with it: launch_mass_cheb)
template <typename T>
void cheb_first(long n, const T* b, const T* P, T* x, T* d, T inv_theta, hipStream_t s);

// Single-workgroup CG on the coarsest level (FP64, identity preconditioner,
// absolute tolerance read from *tol, at most max_its iterations). Work
// vectors g, d, h have the size of x. Writes the iteration count (or -1 on
// no convergence) to *its.
template <int K>
void launch_coarse_cg(const MeshDev& m, int ncomp, bool weighted, const double* rho, double coeff, const double* b,
                      double* x, double* g, double* d, double* h, const double* tol, int max_its, int* its, long Ns,
                      hipStream_t s);

}  // namespace atmo
