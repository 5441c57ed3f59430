// VTU output of a state (rho, u, p), as the reference's output_results
// (Atmospheric_Flow.cc:886-918).
#pragma once
#include <string>
#include <vector>

#include "engine.h"

namespace atmo {

// k: FE_DGQ degree; m: the structured mesh, or cgeo (6 per cell: x0, dx,
// z00, z10, z01, z11) for a locally refined one (then m is not read);
// rho, p: N_s values, u: 2 N_s; high_order: one Lagrange cell of (k+1)^2
// nodes per cell instead of k x k subdivided quadrilaterals.
void write_vtu(int k, const MeshDesc& m, const std::vector<double>& cgeo, const double* rho, const double* u,
               const double* p, const std::string& path, bool high_order);

}  // namespace atmo
