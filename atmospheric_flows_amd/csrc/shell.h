// The reference's own 3D case (Atmospheric_Flow.cc:279-289, 1277): FE_DGQ(1)
// on a spherical shell, MappingQ(2), radial gravity; mesh, geometry and
// operator launchers of the shell path (shell.cpp, shell_ops.hip).
//
// Mesh. The shell between r_in and r_out is the radial extrusion of np
// coarse panels on the sphere (6: the cube's faces, GridGenerator::
// hyper_shell's default coarse mesh in 3D; 12: the rhombi of the rhombic
// dodecahedron), each refined `refinements` times: n = 2^refinements cells
// along each panel edge and across the shell. A panel is the map
// n(s, t) = normalize(bilinear blend of its corner directions d00, d10,
// d01, d11) of (s, t) in [0,1]^2; cell (p, i, j, k) covers
// s in [i, i+1]/n, t in [j, j+1]/n and the radii of layer k. Cells are
// numbered c = ((p*n + k)*n + j)*n + i; dof d = a + 2b + 4c at the cell
// vertex (xi, eta, zeta) = (a, b, c) (FE_DGQ(1), lexicographic); faces
// 0/1: xi = 0/1, 2/3: eta = 0/1, 4/5: zeta = 0/1 (walls at the radii).
//
// Geometry (MappingQ(2) with the support points on the shell): the Q2
// interpolant of the cell map at its 27 support points factorizes as
// X(xi, eta, zeta) = r(zeta) N(xi, eta), with r linear across the layer
// (exact) and N the 2D Q2 interpolant of the panel directions at the 9
// lateral support points; kernels rebuild J from those 9 directions.
//
// Faces. Within a panel and across the shell, neighbours share their face
// coordinates; across a panel edge the coordinate along the edge may run
// the other way: nbc = neighbour face | flip << 3, and face point (gs, gt)
// of this side is (flip ? 2 - gs : gs, gt) of the neighbour's face (face
// coordinates: the two other reference axes in increasing order).
#pragma once
#include <vector>

#include "common.h"

namespace atmo {

struct ShellDesc {
  int refinements = 3;
  int panels = 6;
  double r_inner = 0.9 * 6.37122e6, r_outer = 6.37122e6;  // Atmospheric_Flow.cc:283-284
  double planet = 6.37122e6;                             // a of the gravity factor a^2/r^2
};

struct ShellMesh {
  ShellDesc d;
  int n = 0, ncell = 0;
  std::vector<double> corners;     // [panel][4][3]: d00, d10, d01, d11
  std::vector<int> nb;             // [cell][6]: neighbour cell, -1 at the walls
  std::vector<unsigned char> nbc;  // [cell][6]: neighbour face | flip << 3
};

// Throws std::invalid_argument on a bad description.
ShellMesh build_shell(const ShellDesc& d);
// Physical positions of the dofs (the vertices), [cell][8][3].
void shell_dof_points(const ShellMesh& m, double* out);
// GridTools::minimal_cell_diameter (Atmospheric_Flow.cc:288, 1227): the
// smallest over cells of the longest vertex diagonal.
double shell_min_cell_diameter(const ShellMesh& m);

// The solid-body-rotation steady state of the reference's 3D test
// (equation_data.h:47-176) in its non-dimensional variables: rho = p =
// exp(q), q = (Omega^2 s^2 / 2 - a g (1 - a/r)) / RT with s the distance from
// the polar axis and RT = R T_bar (rho_ref = p0 / (R T_bar), p_ref = p0, so
// both profiles carry the same factor), u = Omega (-y, x, 0).
struct SbrParams {
  double Omega, a, g, RT;
};
// its values at the dofs (VectorTools::interpolate at the mapped support
// points, Atmospheric_Flow.cc:420-422)
void shell_sbr_state(const ShellMesh& m, const SbrParams& sp, double* rho, double* u, double* p);
// analyze_results (Atmospheric_Flow.cc:926-985): L2 errors of a state
// against the steady state with QGauss(k+1) = QGauss(2) on the mapped
// cells; out = {u, u relative, rho, rho relative, p, p relative}
void shell_l2_errors(const ShellMesh& m, const SbrParams& sp, const double* rho, const double* u, const double* p,
                     double* out);

// Device view of a shell mesh.
struct ShellDev {
  int n, ncell;
  double r_in, dr;     // inner radius, layer thickness (r_out - r_in) / n
  double planet2;      // a^2 of the gravity factor a^2 / r^2
  const double* corners;
  const int* nb;
  const unsigned char* nbc;
};

// Physical constants of the shell operators.
struct ShellConsts {
  double Ma, Fr, dt, gamma_gas;
};

// ---- launchers (shell_ops.hip); src lists follow the reference's positional
// conventions (as for the terrain path); velocity fields have 3 components.
void shell_rhs_rho(const ShellDev& m, const ShellConsts& c, int stage, const double* const* src, double* dst,
                   hipStream_t s);
void shell_rhs_pressure(const ShellDev& m, const ShellConsts& c, int stage, const double* const* src, double* dst,
                        hipStream_t s);
void shell_rhs_velocity(const ShellDev& m, const ShellConsts& c, int stage, const double* const* src, double* dst,
                        hipStream_t s);
// B p (vmult_pressure) and C u (vmult_enthalpy, with pfix) for the stage's coefficient
void shell_B(const ShellDev& m, const ShellConsts& c, double coeff, const double* p, double* dst_u, hipStream_t s);
void shell_C(const ShellDev& m, const ShellConsts& c, double coeff, const double* pfix, const double* u,
             double* dst_p, hipStream_t s);
// y = A^{-1} B p with A's packed cell inverse blocks (shell_block_inverse), one pass
void shell_b_ainv(const ShellDev& m, const ShellConsts& c, double coeff, const double* p, const double* inv,
                  double* dst_u, hipStream_t s);
// dst = scale * (D p - C u) and, in the same pass, per-block partial dot products of dst with the columns
// V_0 .. V_{nv-1} of V (leading dimension ldv) and with itself: slot k of block b at part[k * pstride + b],
// slot nv is dst.dst (GMRES's Arnoldi step; when p is column nv-1 its staged values serve that slot)
void shell_d_minus_c_dots(const ShellDev& m, const ShellConsts& c, double coeff, const double* pfix, const double* u,
                          const double* p, const double* scale, double* dst, const double* V, long ldv, int nv,
                          double* part, long pstride, hipStream_t s);
// number of blocks (partials per slot) of shell_d_minus_c_dots
inline int shell_blocks(int ncell) { return (ncell + 6) / 7; }
// dst = scale * (D p - C u) (scale may be null)
void shell_d_minus_c(const ShellDev& m, const ShellConsts& c, double coeff, const double* pfix, const double* u,
                     const double* p, const double* scale, double* dst, hipStream_t s);
// dst = coeff * M src (ncomp components; with rho: the rho-weighted mass A)
void shell_mass(const ShellDev& m, int ncomp, const double* rho, double coeff, const double* src, double* dst,
                hipStream_t s);
// Inverse of every cell block of M (rho == null) or of the rho-weighted
// mass (packed upper triangle, 36 per cell) and the inverse of its diagonal
// (8 per cell); *bad (device, may be null) is set to 1 when a block is not
// positive definite (left unchanged otherwise).
void shell_block_inverse(const ShellDev& m, const double* rho, double* inv, double* inv_diag, double* bad,
                         hipStream_t s);
// x = coeff * Binv b cell by cell (ncomp components)
void shell_block_apply(const ShellDev& m, int ncomp, const double* inv, double coeff, const double* b, double* x,
                       hipStream_t s);

}  // namespace atmo
