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
// matrix_cores: FP32 sweeps at k = 2, 3 on the matrix cores (k_cheb_sweep_mx);
// otherwise, and for every other type and degree, on the vector ALUs.
template <int K, typename T>
void launch_presmooth_residual(const MeshDev& m, int ncomp, bool weighted, T* x, const T* rho, T coeff, const T* b,
                               const T* P, const ChebSteps& cs, T* res, long Ns, bool matrix_cores, hipStream_t s);
// x += prolongation of the coarse level's xc, then one Chebyshev sweep from
// that guess (the V-cycle's coarse-grid correction and post-smoothing).
template <int K, typename T>
void launch_prolongate_postsmooth(const MeshDev& m, const MeshDev& mc, int ncomp, bool weighted, T* x, const T* xc,
                                  const T* rho, T coeff, const T* b, const T* P, const ChebSteps& cs, long Ns,
                                  long Nc, bool matrix_cores, hipStream_t s);

// Single-workgroup CG on the coarsest level (FP64, identity preconditioner,
// absolute tolerance read from *tol, at most max_its iterations). Work
// vectors g, d, h have the size of x. Writes the iteration count (or -1 on
// no convergence) to *its.
template <int K>
void launch_coarse_cg(const MeshDev& m, int ncomp, bool weighted, const double* rho, double coeff, const double* b,
                      double* x, double* g, double* d, double* h, const double* tol, int max_its, int* its, long Ns,
                      hipStream_t s);
// The same for FP32 level vectors b, x (the mixed form): the solve runs in FP64 as above, b and x
// converted as by vec_cast (inside the kernel when the level fits LDS; b64, x64 are the FP64 copies
// used otherwise).
template <int K>
void launch_coarse_cg_f32(const MeshDev& m, int ncomp, bool weighted, const double* rho, double coeff, const float* b,
                          float* x, double* b64, double* x64, double* g, double* d, double* h, const double* tol,
                          int max_its, int* its, long Ns, hipStream_t s);

// Chebyshev eigenvalue estimates of several small levels in one launch: one
// workgroup of 16 waves per level runs all nit CG iterations (Jacobi
// preconditioner P, start vector b, absolute tolerance tol on |g|), with the
// recurrence of the multi-kernel estimate: g = -b, h = P g, d = -h; per
// iteration h = A d, alpha = (g.h)/(d.h) recorded in hist[2 it], x += alpha d,
// g += alpha h, convergence test on |g|, then h = P g, beta = (g.h)_new/(g.h)
// recorded in hist[2 it + 1], d = -h + beta d (the last beta is not formed);
// nothing is recorded once the test has decided. st->it, st->done as the
// multi-kernel estimate leaves them. Work vectors x, g, d, h of the level.
constexpr int EIG_BATCH_MAX = 8;
struct EigLevel {
  MeshDev m;
  const double* rho;  // level density (weighted operator) or null
  const double* P;
  const double* b;
  double *x, *g, *d, *h;
  double* hist;
  CGState* st;
  long Ns;
};
struct EigBatch {
  EigLevel L[EIG_BATCH_MAX];
  int n;
};
template <int K>
void launch_eig_cg_batch(const EigBatch& B, int ncomp, bool weighted, double coeff, int nit, double tol,
                         hipStream_t s);

}  // namespace atmo
