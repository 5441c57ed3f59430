// Per-step diagnostics of EulerSolver::run on device (diag.hip): the
// density bounds printed after each density update and the celerity and
// velocity maxima behind the acoustic and advective Courant numbers
// (Atmospheric_Flow.cc:991-1087, 1121-1122, 1164-1165, 1205-1206, 1225-1235),
// evaluated at the QGaussLobatto<dim>(k+1) points of the reference's FEValues
// (Atmospheric_Flow.cc:86-88, 233-235): FE_DGQ's support points, i.e. the nodal values.
#pragma once
#include "common.h"

namespace atmo {

// Slots of the device-resident step record.
enum DiagSlot {
  DG_MIN_RHO = 0,      // [0..2] min rho at the points after stage 1..3 density update
  DG_MAX_RHO = 3,      // [3..5] max |rho| over the dofs (linfty_norm) of the same fields
  DG_MAX_CEL = 6,      // max sqrt(gamma p / rho) at the points of the new state
  DG_MAX_VEL = 7,      // max |u| at the points of the new state
  DG_NONFINITE = 8,    // [8..11] non-finite values met by each of the four reductions
  DG_SLOTS = 12
};

constexpr int DIAG_MAX_BLOCKS = 1024;
// doubles of the partials buffer one diagnostics call needs
constexpr int DIAG_PARTIALS = 3 * DIAG_MAX_BLOCKS;

// Density record of stage s (0..2): rec[DG_MIN_RHO+s], rec[DG_MAX_RHO+s], rec[DG_NONFINITE+s].
template <int K>
void launch_diag_density(const MeshDev& m, const double* rho, int stage, double* partials, double* rec,
                         hipStream_t s);
// The same for a density field of `items` dofs whose nodal values are the point values (any mesh).
void launch_diag_density_n(long items, const double* rho, int stage, double* partials, double* rec, hipStream_t s);
// Flow record of the state (rho, u, p): rec[DG_MAX_CEL], rec[DG_MAX_VEL], rec[DG_NONFINITE+3].
template <int K>
void launch_diag_flow(const MeshDev& m, const double* rho, const double* u, const double* p, double gamma_gas,
                      long Ns, double* partials, double* rec, hipStream_t s);

// GridTools::minimal_cell_diameter of the terrain mesh (host): min over
// cells of the longer diagonal.
double min_cell_diameter(int nx, int nz, double Lx, double H, const double* h);

}  // namespace atmo
