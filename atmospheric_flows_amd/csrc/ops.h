// Host-side launchers of the operator kernels (ops.hip).
#pragma once
#include "common.h"

namespace atmo {

// Physical/discretization constants shared by the operator kernels.
struct OpConsts {
  double Ma, Fr, dt, gamma_gas;
  double grav[2];     // unit gravity direction e_r (constant for the terrain case)
  double grav_factor; // a^2/r^2 factor of the reference's radial gravity (1 here)
};

// IMEX tableau of euler_operator.h:220-224.
struct Tableau {
  double g, a21, a22, a31, a32, a33, a21t, a22t, a31t, a32t, a33t, b1, b2, b3;
  Tableau();
};

// dst = coeff * M src + add_scale * add  (M: DG mass matrix; NC components)
// WEIGHTED: dst = A src with A = (rho * u, v) (euler_operator.h:2328-2357).
template <int K, typename T>
void launch_mass(const MeshDev& m, int ncomp, bool weighted, const T* src, const T* rho, T* dst, T coeff,
                 const T* add, T add_scale, long Ns, hipStream_t s);

// inverse diagonal of coeff*M (weighted: of A, written to both components)
template <int K, typename T>
void launch_inv_diag(const MeshDev& m, int ncomp, bool weighted, const T* rho, T* dst, T coeff, long Ns,
                     hipStream_t s);

// B (vmult_pressure): p -> u
template <int K>
void launch_B(const MeshDev& m, const OpConsts& oc, double coeff, const double* p, double* dst_u, long Ns,
              hipStream_t s);
// C (vmult_enthalpy): u -> p, with pres_fixed
template <int K>
void launch_C(const MeshDev& m, const OpConsts& oc, double coeff, const double* pfix, const double* u,
              double* dst_p, long Ns, hipStream_t s);
// Schur complement without its inner solve: dst = scale * (D p - C u)
// (scale: optional per-dof factor, e.g. the Jacobi preconditioner of GMRES)
template <int K>
void launch_d_minus_c(const MeshDev& m, const OpConsts& oc, double coeff, const double* pfix, const double* u,
                      const double* p, const double* scale, double* dst, long Ns, hipStream_t s);
// The same product (scale required) and, in the same pass, per-block partial
// dot products of dst with V_0 .. V_{nv-1} (column k at V + k*ldv) and
// dst.dst: slot k of block b at part[k*pstride + b] (slot nv: dst.dst);
// blocks = ceil(ncell / CPW) <= pstride. waves: the waves per SIMD the
// kernel was compiled for (3 or 4, 0: the default for k), results identical.
template <int K>
void launch_d_minus_c_dots(const MeshDev& m, const OpConsts& oc, double coeff, const double* pfix, const double* u,
                           const double* p, const double* scale, double* dst, long Ns, const double* V, long ldv,
                           int nv, double* part, long pstride, int waves, hipStream_t s);
// number of blocks of the cell kernels (the partial count of launch_d_minus_c_dots)
template <int K>
inline int cell_blocks(int ncell) {
  return (ncell + Cfg<K>::CPW - 1) / Cfg<K>::CPW;
}
// explicit residuals; src follows the reference's positional conventions
template <int K>
void launch_rhs_rho(const MeshDev& m, const OpConsts& oc, const Tableau& tb, int stage, const double* const* src,
                    double* dst, long Ns, hipStream_t s);
template <int K>
void launch_rhs_pressure(const MeshDev& m, const OpConsts& oc, const Tableau& tb, int stage,
                         const double* const* src, double* dst, long Ns, hipStream_t s);
template <int K>
void launch_rhs_velocity(const MeshDev& m, const OpConsts& oc, const Tableau& tb, int stage,
                         const double* const* src, double* dst, long Ns, hipStream_t s);

}  // namespace atmo
