// Operator kernels of the spherical-shell path (shell.h): the EULEROperator
// forms of euler_operator.h in dim = 3 at FE_DGQ(1) with QGauss(3)
// (Atmospheric_Flow.cc:345), on MappingQ(2)-mapped cells, radial gravity
// g a^2/r^2 e_r evaluated at every quadrature point.
//
// Launch shape: one wave64 per 7 cells, 9 lanes per cell, one per lateral
// quadrature column. The cell map factorizes as X = r(zeta) N(xi, eta) (the
// shell's Q2 map, shell.h), so J = [r N_xi | r N_eta | dr N]: a lane forms
// the lateral part once (N and the cofactor rows of [N_xi | N_eta | N]) and
// gets the three radial points of its column by scaling with r. The cell's
// dofs are staged through LDS; a neighbour's dofs are read from global
// memory (L2) at the face point its orientation code maps to; no per-point
// geometry is stored.
#include <cmath>
#include <type_traits>

#include "ops.h"
#include "shell.h"
#include "wave.h"

namespace atmo {

namespace {

// QGauss<1>(3) on [0,1]
__constant__ double kGQ[3] = {0.5 - 0.5 * 0.77459666924148337704, 0.5, 0.5 + 0.5 * 0.77459666924148337704};
__constant__ double kGW[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};

__device__ inline double q1v(int i, double x) { return i ? x : 1.0 - x; }
__device__ inline double q1d(int i) { return i ? 1.0 : -1.0; }
// MappingQ(2) Lagrange basis on {0, 1/2, 1}
__device__ inline double q2v(int a, double x) {
  return a == 0 ? 2.0 * (x - 0.5) * (x - 1.0) : (a == 1 ? -4.0 * x * (x - 1.0) : 2.0 * x * (x - 0.5));
}
__device__ inline double q2d(int a, double x) { return a == 0 ? 4.0 * x - 3.0 : (a == 1 ? 4.0 - 8.0 * x : 4.0 * x - 1.0); }

__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// ------------------------------------------------------------------ physics
// Every form below restates its counterpart of euler_operator.h in dim = 3
// (component loops over three velocity components); a Phys type provides the
// NF scalar source fields it reads (velocity fields as three component
// fields), the pointwise cell integrand (value V and flux F tested with
// phi and grad phi), the inner-face flux tested with this side's phi, and
// the wall flux.

struct Weights {
  double wa[3], wt[3];
};

__host__ __device__ inline Weights stage_weights(const Tableau& tb, int st) {
  Weights w{};
  if (st == 1) {
    w.wa[0] = tb.a21;
    w.wt[0] = tb.a21t;
  } else if (st == 2) {
    w.wa[0] = tb.a31;
    w.wt[0] = tb.a31t;
    w.wa[1] = tb.a32;
    w.wt[1] = tb.a32t;
  } else {
    w.wa[0] = tb.b1;
    w.wt[0] = tb.b1;
    w.wa[1] = tb.b2;
    w.wt[1] = tb.b2;
    w.wa[2] = tb.b3;
    w.wt[2] = tb.b3;
  }
  return w;
}

// vmult_rhs_rho_update: cell :344-473, face :476-692, boundary empty (:111-115)
template <int ST>
struct PhysRho {
  static constexpr int NF = 4 * ST, NOUT = 1;
  static constexpr bool FACES = true, GRAVITY = false;
  const double* src[NF];
  double w[3], dt;
  __device__ void cell(const double* v, const double*, double* V, double (*F)[3]) const {
    V[0] = v[0];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double a = 0;
#pragma unroll
      for (int s = 0; s < ST; ++s) a += w[s] * dt * v[4 * s] * v[4 * s + 1 + d];
      F[0][d] = a;
    }
  }
  __device__ void face(const double* vp, const double* vm, const double* n, double* Fv) const {
    double acc = 0;
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      const double rp = vp[4 * s], rm = vm[4 * s];
      const double* up = vp + 4 * s + 1;
      const double* um = vm + 4 * s + 1;
      double fn = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) fn += 0.5 * (rp * up[d] + rm * um[d]) * n[d];
      const double lam = fmax(fabs(dot3(up, n)), fabs(dot3(um, n)));
      acc += w[s] * dt * (fn + 0.5 * lam * (rp - rm));
    }
    Fv[0] = -acc;
  }
  __device__ void wall(const double*, const double*, double* Fv) const { Fv[0] = 0.0; }
};

// vmult_rhs_pressure: cell :745-1007, face :1012-1361, boundary empty.
// Fields: triple s = (rho, u0, u1, u2, p) at 5s, then the kinetic pair
// (rho, u) at 5*ST.
template <int ST>
struct PhysPres {
  static constexpr int NF = 5 * ST + 4, NOUT = 1;
  static constexpr bool FACES = true, GRAVITY = true;
  const double* src[NF];
  Weights w;
  // gravity work terms: weights of the triples 0 .. ST-1 (stage 3) or of the
  // triples 0 .. ST-2 and the kinetic pair (stages 1, 2)
  double gwt[3];
  double dt, Ma2, iFr2, gm, planet2;
  __device__ void cell(const double* v, const double* x, double* V, double (*F)[3]) const {
    const double r = sqrt(dot3(x, x));
    const double er[3] = {x[0] / r, x[1] / r, x[2] / r};
    const double G = planet2 / (r * r);
    double E[ST], uu[ST];
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      const double* u = v + 5 * s + 1;
      uu[s] = dot3(u, u);
      E[s] = 1.0 / (gm - 1.0) * (v[5 * s + 4] / v[5 * s]) + 0.5 * Ma2 * uu[s];
    }
    const double rk = v[5 * ST];
    const double* uk = v + 5 * ST + 1;
    double val = v[0] * E[0] - 0.5 * rk * Ma2 * dot3(uk, uk);
    double gsum = 0;
#pragma unroll
    for (int t = 0; t < ST; ++t) gsum += gwt[t] * v[5 * t] * dot3(v + 5 * t + 1, er);
    if constexpr (ST < 3) gsum += gwt[ST] * rk * dot3(uk, er);
    val -= dt * Ma2 * iFr2 * G * gsum;
    V[0] = val;
    double f[3] = {0, 0, 0};
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      const double rs = v[5 * s];
      const double kin = 0.5 * w.wa[s] * dt * Ma2 * uu[s] * rs;
      const double ent = w.wt[s] * dt * (rs * (E[s] - 0.5 * Ma2 * uu[s]) + v[5 * s + 4]);
#pragma unroll
      for (int d = 0; d < 3; ++d) f[d] += (kin + ent) * v[5 * s + 1 + d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) F[0][d] = f[d];
  }
  __device__ void face(const double* vp, const double* vm, const double* n, double* Fv) const {
    double acc = 0;
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      const double rP = vp[5 * s], rM = vm[5 * s], pP = vp[5 * s + 4], pM = vm[5 * s + 4];
      const double* uP = vp + 5 * s + 1;
      const double* uM = vm + 5 * s + 1;
      const double upp = dot3(uP, uP), umm = dot3(uM, uM);
      const double unP = dot3(uP, n), unM = dot3(uM, n);
      const double akn = 0.5 * (0.5 * upp * rP * unP + 0.5 * umm * rM * unM);
      const double Ep = 1.0 / (gm - 1.0) * pP / rP + 0.5 * Ma2 * upp;
      const double Em = 1.0 / (gm - 1.0) * pM / rM + 0.5 * Ma2 * umm;
      const double hp = (Ep - 0.5 * Ma2 * upp) * rP + pP;
      const double hm = (Em - 0.5 * Ma2 * umm) * rM + pM;
      const double aen = 0.5 * (hp * unP + hm * unM);
      const double lam = fmax(fabs(unP), fabs(unM));
      if constexpr (ST < 3) {
        double jkin;
        if (ST == 2 && s == 1)  // euler_operator.h:1182-1183 (u_tmp_2 . u_old, as the reference writes it)
          jkin = rP * 0.5 * dot3(uP, vp + 1) - rM * 0.5 * dot3(uM, vm + 1);
        else
          jkin = rP * 0.5 * upp - rM * 0.5 * umm;
        const double je = rP * (Ep - 0.5 * Ma2 * upp) - rM * (Em - 0.5 * Ma2 * umm);
        acc += -w.wa[s] * dt * Ma2 * (akn + 0.5 * lam * jkin) - w.wt[s] * dt * (aen + 0.5 * lam * je);
      } else {
        const double jrE = rP * Ep - rM * Em;
        acc += -w.wa[s] * dt * Ma2 * akn - w.wa[s] * dt * aen - w.wa[s] * dt * 0.5 * lam * jrE;
      }
    }
    Fv[0] = acc;
  }
  __device__ void wall(const double*, const double*, double* Fv) const { Fv[0] = 0.0; }
};

// vmult_rhs_velocity_fixed: cell :1620-1886, face :1889-2184, boundary
// :2187-2305. Fields: triple s at 5s, then (ST < 3) the density of the
// implicit stage's gravity term at 5*ST.
template <int ST>
struct PhysVel {
  static constexpr int NF = 5 * ST + (ST < 3 ? 1 : 0), NOUT = 3;
  static constexpr bool FACES = true, GRAVITY = true;
  const double* src[NF];
  Weights w;
  double gw[3];
  double dt, Ma2, iFr2, planet2;
  __device__ void cell(const double* v, const double* x, double* V, double (*F)[3]) const {
    const double r = sqrt(dot3(x, x));
    const double er[3] = {x[0] / r, x[1] / r, x[2] / r};
    const double G = planet2 / (r * r);
    // gravity densities: triples 0 and 1 (stage >= 2) and the extra field or triple 2
    double gsum = gw[0] * v[0] + gw[1] * v[5];
    if constexpr (ST >= 2) gsum += gw[2] * v[10];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      V[c] = v[0] * v[1 + c] - dt * iFr2 * gsum * er[c] * G;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double a = 0;
#pragma unroll
        for (int s = 0; s < ST; ++s)
          a += w.wa[s] * dt * v[5 * s] * v[5 * s + 1 + c] * v[5 * s + 1 + d] + (d == c ? w.wt[s] * dt / Ma2 * v[5 * s + 4] : 0.0);
        F[c][d] = a;
      }
    }
  }
  __device__ void face(const double* vp, const double* vm, const double* n, double* Fv) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) Fv[c] = 0;
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      const double rP = vp[5 * s], rM = vm[5 * s];
      const double* uP = vp + 5 * s + 1;
      const double* uM = vm + 5 * s + 1;
      const double unP = dot3(uP, n), unM = dot3(uM, n);
      const double avgp = 0.5 * (vp[5 * s + 4] + vm[5 * s + 4]);
      const double lam = fmax(fabs(unP), fabs(unM));
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double tn = 0.5 * (rP * uP[c] * unP + rM * uM[c] * unM);
        const double jr = rP * uP[c] - rM * uM[c];
        Fv[c] += -w.wa[s] * dt * tn - w.wt[s] * dt / Ma2 * avgp * n[c] - w.wa[s] * dt * 0.5 * lam * jr;
      }
    }
  }
  __device__ void wall(const double* vp, const double* n, double* Fv) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) Fv[c] = 0;
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      const double pP = vp[5 * s + 4], pD = pP;
#pragma unroll
      for (int c = 0; c < 3; ++c) Fv[c] += -w.wt[s] * dt / Ma2 * 0.5 * (pP + pD) * n[c];
    }
  }
};

// vmult_pressure (B): cell :1510-1560, face :1563-1598, boundary :1601-1616
struct PhysB {
  static constexpr int NF = 1, NOUT = 3;
  static constexpr bool FACES = true, GRAVITY = false;
  const double* src[NF];
  double cf;
  __device__ void cell(const double* v, const double*, double* V, double (*F)[3]) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      V[c] = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) F[c][d] = d == c ? -cf * v[0] : 0.0;
    }
  }
  __device__ void face(const double* vp, const double* vm, const double* n, double* Fv) const {
    const double avg = 0.5 * (vp[0] + vm[0]);
#pragma unroll
    for (int c = 0; c < 3; ++c) Fv[c] = cf * avg * n[c];
  }
  __device__ void wall(const double* vp, const double* n, double* Fv) const {
    const double pD = vp[0];
#pragma unroll
    for (int c = 0; c < 3; ++c) Fv[c] = cf * 0.5 * (vp[0] + pD) * n[c];
  }
};

// vmult_enthalpy (C): cell :1411-1460, face :1463-1507, boundary empty.
// Fields (p_fix, u0, u1, u2) at OFF; with DMC also D p (p at field 0):
// the Schur complement's D - C of apply_add NS 2 (:2374-2408).
template <bool DMC>
struct PhysC {
  static constexpr int OFF = DMC ? 1 : 0;
  static constexpr int NF = 4 + OFF, NOUT = 1;
  static constexpr bool FACES = true, GRAVITY = false;
  const double* src[NF];
  double coeff, dt, gm;
  __device__ void cell(const double* v, const double*, double* V, double (*F)[3]) const {
    const double s = -coeff * dt * gm / (gm - 1.0) * v[OFF];
    const double sg = DMC ? -1.0 : 1.0;
    V[0] = DMC ? v[0] / (gm - 1.0) : 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) F[0][d] = sg * s * v[OFF + 1 + d];
  }
  __device__ void face(const double* vp, const double* vm, const double* n, double* Fv) const {
    const double* uP = vp + OFF + 1;
    const double* uM = vm + OFF + 1;
    const double unP = dot3(uP, n), unM = dot3(uM, n);
    const double an = 0.5 * gm / (gm - 1.0) * (vp[OFF] * unP + vm[OFF] * unM);
    const double lam = fmax(fabs(unP), fabs(unM));
    const double jump = 1.0 / (gm - 1.0) * (vp[OFF] - vm[OFF]);
    const double F = coeff * dt * (an + 0.5 * lam * jump);
    Fv[0] = DMC ? -F : F;
  }
  __device__ void wall(const double*, const double*, double* Fv) const { Fv[0] = 0.0; }
};

// mass-type blocks: coeff * M src (NC components), rho-weighted when W
// (euler_operator.h:717-740, 2328-2357)
template <int NC, bool W>
struct PhysMass {
  static constexpr int NF = NC + (W ? 1 : 0), NOUT = NC;
  static constexpr bool FACES = false, GRAVITY = false;
  const double* src[NF];
  double coeff;
  __device__ void cell(const double* v, const double*, double* V, double (*F)[3]) const {
    const double wq = W ? coeff * v[NC] : coeff;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      V[c] = wq * v[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) F[c][d] = 0.0;
    }
  }
  __device__ void face(const double*, const double*, const double*, double*) const {}
  __device__ void wall(const double*, const double*, double*) const {}
};

// ------------------------------------------------------------------ kernel
// Lateral support direction t = a + 3b (a, b in {0, 1/2, 1}) of the cell's panel map.
__device__ inline void support_dir(const ShellDev& m, int cell, int t, double* out) {
  const int n = m.n;
  const int i = cell % n, j = (cell / n) % n, p = cell / (n * n * n);
  const int a = t % 3, b = t / 3;
  const double s = (i + 0.5 * a) / n, u = (j + 0.5 * b) / n;
  const double* C = m.corners + 12 * p;
  double v[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    v[d] = (1 - s) * (1 - u) * C[d] + s * (1 - u) * C[3 + d] + (1 - s) * u * C[6 + d] + s * u * C[9 + d];
  const double nn = sqrt(dot3(v, v));
#pragma unroll
  for (int d = 0; d < 3; ++d) out[d] = v[d] / nn;
}

// Lateral geometry of the map X = r(zeta) N(xi, eta) at (xi, eta): N and the
// cofactor rows of A = [N_xi | N_eta | N]: C0 = N_eta x N, C1 = N x N_xi,
// C2 = N_xi x N_eta, det A = N_xi . C0. With J = A diag(r, r, dr):
// det J = r^2 dr det A, rows of det J * J^{-1}: r dr C0, r dr C1, r^2 C2.
struct LatGeo {
  double N[3], C0[3], C1[3], C2[3], det;
};

__device__ inline void lat_geo(const double (*S)[3], double xi, double eta, LatGeo& g) {
  double La[3], Da[3], Lb[3], Db[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    La[a] = q2v(a, xi);
    Da[a] = q2d(a, xi);
    Lb[a] = q2v(a, eta);
    Db[a] = q2d(a, eta);
  }
  double N0 = 0, N1 = 0, N2 = 0, X0 = 0, X1 = 0, X2 = 0, Y0 = 0, Y1 = 0, Y2 = 0;
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double w = La[a] * Lb[b], wx = Da[a] * Lb[b], wy = La[a] * Db[b];
      const double v0 = S[a + 3 * b][0], v1 = S[a + 3 * b][1], v2 = S[a + 3 * b][2];
      N0 += w * v0;
      N1 += w * v1;
      N2 += w * v2;
      X0 += wx * v0;
      X1 += wx * v1;
      X2 += wx * v2;
      Y0 += wy * v0;
      Y1 += wy * v1;
      Y2 += wy * v2;
    }
  g.N[0] = N0;
  g.N[1] = N1;
  g.N[2] = N2;
  g.C0[0] = Y1 * N2 - Y2 * N1;
  g.C0[1] = Y2 * N0 - Y0 * N2;
  g.C0[2] = Y0 * N1 - Y1 * N0;
  g.C1[0] = N1 * X2 - N2 * X1;
  g.C1[1] = N2 * X0 - N0 * X2;
  g.C1[2] = N0 * X1 - N1 * X0;
  g.C2[0] = X1 * Y2 - X2 * Y1;
  g.C2[1] = X2 * Y0 - X0 * Y2;
  g.C2[2] = X0 * Y1 - X1 * Y0;
  g.det = X0 * g.C0[0] + X1 * g.C0[1] + X2 * g.C0[2];
}

// bilinear values at (xi, eta) of the bottom (zeta = 0) and top (zeta = 1)
// dofs of a Q1 cell field u[8]
__device__ inline void lat_vals(const double* u, double xi, double eta, double& vb, double& vt) {
  const double w0 = (1 - xi) * (1 - eta), w1 = xi * (1 - eta), w2 = (1 - xi) * eta, w3 = xi * eta;
  vb = u[0] * w0 + u[1] * w1 + u[2] * w2 + u[3] * w3;
  vt = u[4] * w0 + u[5] * w1 + u[6] * w2 + u[7] * w3;
}

// Packed upper triangle of a symmetric 8 x 8 block: entry (a, b), a <= b.
__host__ __device__ constexpr int sym8(int a, int b) { return a <= b ? a * (15 - a) / 2 + b : b * (15 - b) / 2 + a; }

// Cells per wave: 7 cells x 9 lanes (one per lateral quadrature column (a, b)).
constexpr int SCPW = 7;

// Lane (cell w, column t = a + 3b) forms the lateral geometry of its column
// once and evaluates the cell integrand at the column's three radial
// quadrature points, then the bottom and top face points of the column;
// the 12 (lateral face, edge point) pairs of the cell go to lanes t and
// t - 9, each with its three radial points. Face fluxes get the area vector
// (unit normal times JxW) in place of the normal: every flux is positively
// homogeneous of degree 1 in the normal. Each lane accumulates its points'
// contributions to all 8 NOUT (test function, component) integrals in
// registers; the nine lanes of a cell are then summed through LDS.
// What a k_shell launch does with its result (EPI): 0 stores it (times
// scale when given); 1 stores it through the cells' inverse blocks (packed,
// 36 per cell: B p -> A^{-1} B p in one pass); 2 stores it times scale and,
// in the same pass, forms the per-block partial dot products of the result
// with the basis vectors V_0 .. V_{nv-1} and with itself (GMRES's Arnoldi
// step; NOUT = 1).
struct ShellEpi {
  const double* inv = nullptr;
  const double* V = nullptr;
  long ldv = 0;
  int nv = 0, p_last = 0;
  double* part = nullptr;
  long pstride = 0;
};

// Register budget: the Schur complement's kernels (C, D - C, B through the inverse blocks) are compiled
// for 3 waves per SIMD (168 VGPRs, 8-16 bytes of spills per lane; measured faster than 2 waves at 182-206
// VGPRs); the residual kernels keep the compiler's choice (at 168 VGPRs they spill hundreds of bytes).
template <class Phys, int EPI>
constexpr int shell_waves() {
  if constexpr (std::is_same_v<Phys, PhysB>)
    return EPI == 1 ? 3 : 1;
  else if constexpr (std::is_same_v<Phys, PhysC<false>> || std::is_same_v<Phys, PhysC<true>>)
    return 3;
  else
    return 1;
}

template <class Phys, int EPI>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(shell_waves<Phys, EPI>())))
k_shell(ShellDev m, Phys ph, double* __restrict__ dst, long Ns, const double* __restrict__ scale, ShellEpi ep) {
  constexpr int NF = Phys::NF, NO = Phys::NOUT, CPW = SCPW, NI = 8 * NO;
  static_assert(EPI != 2 || NO == 1, "Arnoldi epilogue: one output field");
  __shared__ double sdof[NF][CPW * 8];
  __shared__ double sS[CPW][9][3];
  __shared__ double sred[CPW][9][NI];
  __shared__ double sinv[EPI == 1 ? CPW * 36 : 1];   // EPI 1: the wave's inverse blocks
  __shared__ double sbp[EPI == 1 ? CPW : 1][NI];     // EPI 1: the cells' results before the inverse
  const int lane = threadIdx.x;
  const int w = lane / 9, t = lane - 9 * w;
  const int c0 = cell_block() * CPW;
  const int cell = c0 + w;
  const bool active = w < CPW && cell < m.ncell;
  const long lim = (long)m.ncell * 8;
  // neighbour indices and orientation codes of the faces this lane evaluates (bottom, top and its one or
  // two lateral faces), loaded with the staged fields so they share that memory latency
  const int pr0 = t, pr1 = t + 9;  // the lane's (lateral face, edge point) pairs; pr1 only when < 12
  int nbz[2] = {-1, -1}, nbl[2] = {-1, -1};
  unsigned char cbz[2] = {0, 0}, cbl[2] = {0, 0};
  if (Phys::FACES && active) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      nbz[k] = m.nb[(long)cell * 6 + 4 + k];
      cbz[k] = m.nbc[(long)cell * 6 + 4 + k];
    }
    nbl[0] = m.nb[(long)cell * 6 + pr0 / 3];
    cbl[0] = m.nbc[(long)cell * 6 + pr0 / 3];
    if (pr1 < 12) {
      nbl[1] = m.nb[(long)cell * 6 + pr1 / 3];
      cbl[1] = m.nbc[(long)cell * 6 + pr1 / 3];
    }
  }
  // EPI 1: the wave's inverse blocks (independent of the result), loaded with the staged fields
  constexpr int NIV = (CPW * 36 + 63) / 64;
  double ivr[EPI == 1 ? NIV : 1];
  if constexpr (EPI == 1) {
#pragma unroll
    for (int k = 0; k < NIV; ++k) {
      const int q = lane + 64 * k;
      ivr[k] = (q < CPW * 36 && c0 + q / 36 < m.ncell) ? ep.inv[(long)c0 * 36 + q] : 0.0;
    }
  }
  // staged fields: fixed trip counts, every load of a group of passes issued before its first LDS write
  {
    constexpr int NST = NF * CPW * 8, NSP = (NST + 63) / 64, G = 8;
#pragma unroll
    for (int g0 = 0; g0 < NSP; g0 += G) {
      double sv[G];
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const int q = lane + 64 * (g0 + k);
        const int f = q / (CPW * 8), r = q - f * (CPW * 8);
        const long g = (long)c0 * 8 + r;
        sv[k] = (g0 + k < NSP && q < NST && g < lim) ? ph.src[f < NF ? f : 0][g] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const int q = lane + 64 * (g0 + k);
        const int f = q / (CPW * 8), r = q - f * (CPW * 8);
        if (g0 + k < NSP && q < NST) sdof[f][r] = sv[k];
      }
    }
  }
  if constexpr (EPI == 1) {
#pragma unroll
    for (int k = 0; k < NIV; ++k) {
      const int q = lane + 64 * k;
      if (q < CPW * 36) sinv[q] = ivr[k];
    }
  }
  double r0 = 0;
  if (active) {
    support_dir(m, cell, t, sS[w][t]);
    r0 = m.r_in + m.dr * ((cell / (m.n * m.n)) % m.n);
  }
  __syncthreads();
  const double dr = m.dr;
  double acc[NI];
#pragma unroll
  for (int k = 0; k < NI; ++k) acc[k] = 0.0;
  if (active) {
    const int a = t % 3, b = t / 3;
    const double la[2] = {1.0 - kGQ[a], kGQ[a]}, lb[2] = {1.0 - kGQ[b], kGQ[b]};
    LatGeo G;
    lat_geo(sS[w], kGQ[a], kGQ[b], G);
    double vb[NF], vt[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) lat_vals(&sdof[f][w * 8], kGQ[a], kGQ[b], vb[f], vt[f]);
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
      const double z = kGQ[c], r = r0 + dr * z, w3 = kGW[a] * kGW[b] * kGW[c];
      const double lc[2] = {1.0 - z, z};
      double v[NF], x[3], V[NO], F[NO][3];
#pragma unroll
      for (int f = 0; f < NF; ++f) v[f] = vb[f] + z * (vt[f] - vb[f]);
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] = r * G.N[d];
      ph.cell(v, x, V, F);
      const double JxW = r * r * dr * G.det * w3, s01 = r * dr * w3, s2 = r * r * w3;
#pragma unroll
      for (int o = 0; o < NO; ++o) {
        const double v0 = V[o] * JxW, g0 = dot3(G.C0, F[o]) * s01, g1 = dot3(G.C1, F[o]) * s01,
                     g2 = dot3(G.C2, F[o]) * s2;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int ia = i & 1, ib = (i >> 1) & 1, ic = i >> 2;
          const double sa = ia ? 1.0 : -1.0, sb = ib ? 1.0 : -1.0, sc = ic ? 1.0 : -1.0;
          acc[8 * o + i] += v0 * la[ia] * lb[ib] * lc[ic] + g0 * sa * lb[ib] * lc[ic] + g1 * la[ia] * sb * lc[ic] +
                            g2 * la[ia] * lb[ib] * sc;
        }
      }
    }
    if (Phys::FACES) {
      // bottom (4) and top (5) faces of the column (unrolled: `own` names one register array)
      // the neighbours' traces at this column's points of both faces, their loads issued together
      double vmz[2][NF];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int nbr = nbz[k] < 0 ? cell : nbz[k];  // a wall reads its own cell (the value is not used)
        const int g = cbz[k] & 7, a2 = (cbz[k] >> 3) ? 2 - a : a;
        const double xi = kGQ[a2], eta = kGQ[b];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const double* u = ph.src[f] + (long)nbr * 8 + (g == 5 ? 4 : 0);
          vmz[k][f] = u[0] * (1 - xi) * (1 - eta) + u[1] * xi * (1 - eta) + u[2] * (1 - xi) * eta + u[3] * xi * eta;
        }
      }
#pragma unroll
      for (int fz = 4; fz < 6; ++fz) {
        const double rr = fz == 4 ? r0 : r0 + dr, sg = fz == 4 ? -1.0 : 1.0;
        const double ws = sg * rr * rr * kGW[a] * kGW[b];
        const double Av[3] = {ws * G.C2[0], ws * G.C2[1], ws * G.C2[2]};
        const double* own = fz == 4 ? vb : vt;
        double Fv[NO];
        if (nbz[fz - 4] < 0) {
          ph.wall(own, Av, Fv);
        } else {
          ph.face(own, vmz[fz - 4], Av, Fv);
        }
#pragma unroll
        for (int o = 0; o < NO; ++o)
#pragma unroll
          for (int i = (fz - 4) * 4; i < (fz - 4) * 4 + 4; ++i)  // the test functions with ic = fz - 4
            acc[8 * o + i] += Fv[o] * la[i & 1] * lb[(i >> 1) & 1];
      }
      // lateral faces: pairs (f, e) of face f < 4 and edge point e
#pragma unroll 1
      for (int pr = pr0; pr < 12; pr += 9) {
        const int f = pr / 3, e = pr % 3;
        const double xi = f < 2 ? (double)(f & 1) : kGQ[e], eta = f < 2 ? kGQ[e] : (double)(f & 1);
        const double sg = (f & 1) ? 1.0 : -1.0;
        // area direction on the face: C0 = N_eta x N (xi faces) or C1 = N x N_xi (eta faces). On a face
        // one Lagrange factor of the map is a unit vector, so N and its derivative along the face's edge
        // are sums over the three support directions of that edge alone: the same nonzero terms, in the
        // same order, as lat_geo's full sums, whose other terms are exact zeros
        double Nf[3] = {0.0, 0.0, 0.0}, Nd[3] = {0.0, 0.0, 0.0};
        {
          const int s0 = f < 2 ? 2 * (f & 1) : 6 * (f & 1), st = f < 2 ? 3 : 1;
          const double ue = kGQ[e];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double Lk = q2v(k, ue), Dk = q2d(k, ue);
            const double* T = sS[w][s0 + st * k];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              Nf[d] += Lk * T[d];
              Nd[d] += Dk * T[d];
            }
          }
        }
        // blended arithmetically: a per-lane select between the two arrays becomes a select of addresses
        const double k0 = f < 2 ? 1.0 : 0.0, k1 = 1.0 - k0;
        const double C0[3] = {Nd[1] * Nf[2] - Nd[2] * Nf[1], Nd[2] * Nf[0] - Nd[0] * Nf[2],
                              Nd[0] * Nf[1] - Nd[1] * Nf[0]};
        const double C1[3] = {Nf[1] * Nd[2] - Nf[2] * Nd[1], Nf[2] * Nd[0] - Nf[0] * Nd[2],
                              Nf[0] * Nd[1] - Nf[1] * Nd[0]};
        double Cf[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) Cf[d] = k0 * C0[d] + k1 * C1[d];
        // test functions on face f, by value: the lateral factor along the face's edge
        // coordinate, zero off the face
        double wl[4];  // i & 3 -> lateral weight of test function (i & 1, (i >> 1) & 1)
        const double le[2] = {1.0 - kGQ[e], kGQ[e]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ia = q & 1, ib = q >> 1;
          wl[q] = f < 2 ? (ia == (f & 1) ? le[ib] : 0.0) : (ib == (f & 1) ? le[ia] : 0.0);
        }
        double ob[NF], ot[NF], nb_[NF], nt[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) lat_vals(&sdof[q][w * 8], xi, eta, ob[q], ot[q]);
        const int nbr = pr == pr0 ? nbl[0] : nbl[1];
        if (nbr >= 0) {
          // the neighbour's trace along its face g at edge point e2: its two bottom and
          // two top dofs on that face, linear in the coordinate along the edge
          const int code = pr == pr0 ? cbl[0] : cbl[1];
          const int g = code & 7, e2 = (code >> 3) ? 2 - e : e;
          const int i0 = g < 2 ? (g & 1) : 2 * (g & 1), step = g < 2 ? 2 : 1;
          const double y = kGQ[e2];
#pragma unroll
          for (int q = 0; q < NF; ++q) {
            const double* u = ph.src[q] + (long)nbr * 8 + i0;
            nb_[q] = (1 - y) * u[0] + y * u[step];
            nt[q] = (1 - y) * u[4] + y * u[4 + step];
          }
        }
#pragma unroll 1
        for (int gt = 0; gt < 3; ++gt) {
          const double z = kGQ[gt], r = r0 + dr * z;
          const double ws = sg * r * dr * kGW[e] * kGW[gt];
          const double Av[3] = {ws * Cf[0], ws * Cf[1], ws * Cf[2]};
          double vp[NF], Fv[NO];
#pragma unroll
          for (int q = 0; q < NF; ++q) vp[q] = ob[q] + z * (ot[q] - ob[q]);
          if (nbr < 0) {
            ph.wall(vp, Av, Fv);
          } else {
            double vm[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q) vm[q] = nb_[q] + z * (nt[q] - nb_[q]);
            ph.face(vp, vm, Av, Fv);
          }
          const double lz[2] = {1.0 - z, z};
#pragma unroll
          for (int o = 0; o < NO; ++o)
#pragma unroll
            for (int i = 0; i < 8; ++i) acc[8 * o + i] += Fv[o] * wl[i & 3] * lz[i >> 2];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NI; ++k) sred[w][t][k] = acc[k];
  }
  __syncthreads();
  if constexpr (EPI == 0) {
    if (active) {
#pragma unroll 1
      for (int it = t; it < NI; it += 9) {
        double sum = 0.0;
#pragma unroll
        for (int l = 0; l < 9; ++l) sum += sred[w][l][it];
        const long k = (long)(it >> 3) * Ns + (long)cell * 8 + (it & 7);
        dst[k] = scale ? sum * scale[k] : sum;
      }
    }
  } else if constexpr (EPI == 1) {
    // dof a of output o: row a of the cell's inverse block times the cell's result of that output (the
    // expression of k_shell_block_apply with coeff 1)
    if (active) {
#pragma unroll 1
      for (int it = t; it < NI; it += 9) {
        double sum = 0.0;
#pragma unroll
        for (int l = 0; l < 9; ++l) sum += sred[w][l][it];
        sbp[w][it] = sum;
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll 1
      for (int it = t; it < NI; it += 9) {
        const int o = it >> 3, a = it & 7;
        double acc = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) acc += sinv[w * 36 + sym8(a, j)] * sbp[w][8 * o + j];
        dst[(long)o * Ns + (long)cell * 8 + a] = 1.0 * acc;
      }
    }
  } else {
    // lane t < 8 of a cell holds dof t of the result; the dot products with the basis go in chunks of 8
    // slots (slot k: column k of V) whose loads are all in flight at once, then w.p (with p_last: p is
    // the staged field 0, column nv-1 of V) and w.w
    const bool own = active && t < 8;
    const long k = (long)cell * 8 + t;
    double wv = 0.0;
    if (own) {
      double sum = 0.0;
#pragma unroll
      for (int l = 0; l < 9; ++l) sum += sred[w][l][t];
      wv = scale ? sum * scale[k] : sum;
      dst[k] = wv;
    }
    const long gk = own ? k : 0;
    const int nload = ep.p_last ? ep.nv - 1 : ep.nv;
    const int blk = cell_block();
    const int l = lane & 63;
    for (int k0 = 0; k0 < nload; k0 += 8) {
      double v[8];
#pragma unroll
      for (int s8 = 0; s8 < 8; ++s8) {
        const bool ld = k0 + s8 < nload;  // uniform
        const double y = ld ? __builtin_nontemporal_load(ep.V + (long)(k0 + s8) * ep.ldv + gk) : 0.0;
        v[s8] = wv * y;
      }
      const double sum = wave_sum8(v);
      const int slot = k0 + (l >> 3);
      if ((l & 7) == 0 && slot < nload) ep.part[(long)slot * ep.pstride + blk] = sum;
    }
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (ep.p_last && own) v[0] = wv * sdof[0][w * 8 + t];
    v[1] = wv * wv;
    const double sum = wave_sum8(v);
    if (l == 0 && ep.p_last) ep.part[(long)(ep.nv - 1) * ep.pstride + blk] = sum;
    if (l == 8) ep.part[(long)ep.nv * ep.pstride + blk] = sum;
  }
}

template <class Phys, int EPI = 0>
void launch(const ShellDev& m, const Phys& ph, double* dst, const double* scale, hipStream_t s,
            const ShellEpi& ep = ShellEpi()) {
  const int blocks = (m.ncell + SCPW - 1) / SCPW;
  hipLaunchKernelGGL((k_shell<Phys, EPI>), dim3(blocks), dim3(64), 0, s, m, ph, dst, (long)m.ncell * 8, scale, ep);
  ATMO_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ cell blocks

// Inverse of each cell's mass block (weighted by rho when given), by
// Cholesky in one lane per cell; the inverse diagonal of the block itself too.
__global__ void __launch_bounds__(64) k_shell_blocks(ShellDev m, const double* __restrict__ rho,
                                                     double* __restrict__ inv, double* __restrict__ inv_diag,
                                                     double* __restrict__ bad) {
  constexpr int CPW = SCPW;
  __shared__ double sS[CPW][9][3];
  __shared__ double swq[CPW][27];
  __shared__ double sB[CPW][36];
  const int lane = threadIdx.x;
  const int w = lane / 9, t = lane - 9 * w;
  const int cell = blockIdx.x * CPW + w;
  const bool active = w < CPW && cell < m.ncell;
  double r0 = 0;
  if (active) {
    support_dir(m, cell, t, sS[w][t]);
    r0 = m.r_in + m.dr * ((cell / (m.n * m.n)) % m.n);
  }
  __syncthreads();
  if (active) {
    const int a = t % 3, b = t / 3;
    LatGeo G;
    lat_geo(sS[w], kGQ[a], kGQ[b], G);
    double rb = 1.0, rt = 1.0;
    if (rho) lat_vals(rho + (long)cell * 8, kGQ[a], kGQ[b], rb, rt);
    for (int c = 0; c < 3; ++c) {
      const double z = kGQ[c], r = r0 + m.dr * z;
      const double wq = r * r * m.dr * G.det * kGW[a] * kGW[b] * kGW[c];
      swq[w][t + 9 * c] = rho ? wq * (rb + z * (rt - rb)) : wq;
    }
  }
  __syncthreads();
  if (active) {
    for (int e = t; e < 36; e += 9) {
      int a = 0;
      while (sym8(a, 7) < e) ++a;
      const int b = e - sym8(a, a) + a;
      double acc = 0;
      for (int q = 0; q < 27; ++q) {
        const double xi[3] = {kGQ[q % 3], kGQ[(q / 3) % 3], kGQ[q / 9]};
        const double pa = q1v(a & 1, xi[0]) * q1v((a >> 1) & 1, xi[1]) * q1v(a >> 2, xi[2]);
        const double pb = q1v(b & 1, xi[0]) * q1v((b >> 1) & 1, xi[1]) * q1v(b >> 2, xi[2]);
        acc += swq[w][q] * pa * pb;
      }
      sB[w][e] = acc;
    }
  }
  __syncthreads();
  if (active && t == 0) {
    double L[8][8];
    bool spd = true;
    for (int j = 0; j < 8; ++j) {
      double s = sB[w][sym8(j, j)];
      for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
      spd = spd && s > 0.0;  // fails for a density that is not positive in the cell (or NaN)
      L[j][j] = sqrt(s);
      for (int i = j + 1; i < 8; ++i) {
        double u = sB[w][sym8(j, i)];
        for (int k = 0; k < j; ++k) u -= L[i][k] * L[j][k];
        L[i][j] = u / L[j][j];
      }
    }
    // Linv (lower), then inv = Linv^T Linv
    double Li[8][8];
    for (int c = 0; c < 8; ++c)
      for (int i = 0; i < 8; ++i) {
        if (i < c) {
          Li[i][c] = 0.0;
          continue;
        }
        double u = i == c ? 1.0 : 0.0;
        for (int k = c; k < i; ++k) u -= L[i][k] * Li[k][c];
        Li[i][c] = u / L[i][i];
      }
    for (int a = 0; a < 8; ++a)
      for (int b = a; b < 8; ++b) {
        double u = 0;
        for (int k = b; k < 8; ++k) u += Li[k][a] * Li[k][b];
        inv[(long)cell * 36 + sym8(a, b)] = u;
      }
    for (int a = 0; a < 8; ++a) inv_diag[(long)cell * 8 + a] = 1.0 / sB[w][sym8(a, a)];
    if (!spd && bad) *bad = 1.0;
  }
}

// x = coeff * Binv b, one thread per (component, cell, row)
__global__ void k_shell_block_apply(int ncell, int ncomp, const double* __restrict__ inv, double coeff,
                                    const double* __restrict__ b, double* __restrict__ x) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long Ns = (long)ncell * 8;
  if (t >= Ns * ncomp) return;
  const long comp = t / Ns, r = t - comp * Ns, cell = r >> 3;
  const int a = (int)(r & 7);
  const double* B = inv + cell * 36;
  const double* bb = b + comp * Ns + cell * 8;
  double acc = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) acc += B[sym8(a, j)] * bb[j];
  x[t] = coeff * acc;
}

}  // namespace

// ------------------------------------------------------------ launchers
static void velocity_fields(const double* u, long Ns, const double** out) {
  out[0] = u;
  out[1] = u + Ns;
  out[2] = u + 2 * Ns;
}

void shell_rhs_rho(const ShellDev& m, const ShellConsts& c, int stage, const double* const* src, double* dst,
                   hipStream_t s) {
  const long Ns = (long)m.ncell * 8;
  const Tableau tb;
  double w[3] = {0, 0, 0};
  if (stage == 1) {
    w[0] = tb.a21;
  } else if (stage == 2) {
    w[0] = tb.a31;
    w[1] = tb.a32;
  } else {
    w[0] = tb.b1;
    w[1] = tb.b2;
    w[2] = tb.b3;
  }
  auto fill = [&](auto& ph, int ns_) {
    for (int q = 0; q < ns_; ++q) {
      ph.src[4 * q] = src[2 * q];
      velocity_fields(src[2 * q + 1], Ns, &ph.src[4 * q + 1]);
    }
    for (int q = 0; q < 3; ++q) ph.w[q] = w[q];
    ph.dt = c.dt;
  };
  if (stage == 1) {
    PhysRho<1> ph;
    fill(ph, 1);
    launch(m, ph, dst, nullptr, s);
  } else if (stage == 2) {
    PhysRho<2> ph;
    fill(ph, 2);
    launch(m, ph, dst, nullptr, s);
  } else {
    PhysRho<3> ph;
    fill(ph, 3);
    launch(m, ph, dst, nullptr, s);
  }
}

void shell_rhs_pressure(const ShellDev& m, const ShellConsts& c, int stage, const double* const* src, double* dst,
                        hipStream_t s) {
  const long Ns = (long)m.ncell * 8;
  const Tableau tb;
  auto fill = [&](auto& ph, int st) {
    for (int q = 0; q < st; ++q) {
      ph.src[5 * q] = src[3 * q];
      velocity_fields(src[3 * q + 1], Ns, &ph.src[5 * q + 1]);
      ph.src[5 * q + 4] = src[3 * q + 2];
    }
    ph.src[5 * st] = src[3 * st];
    velocity_fields(src[3 * st + 1], Ns, &ph.src[5 * st + 1]);
    ph.w = stage_weights(tb, st);
    if (st == 1) {  // triple 0, then the kinetic pair (rho_tmp_2, u_fixed)
      ph.gwt[0] = tb.a21t;
      ph.gwt[1] = tb.a22t;
    } else if (st == 2) {  // triples 0 and 1, then the kinetic pair (rho_tmp_3, u_fixed)
      ph.gwt[0] = tb.a31t;
      ph.gwt[1] = tb.a32t;
      ph.gwt[2] = tb.a33t;
    } else {  // triples 0, 1 and 2
      ph.gwt[0] = tb.b1;
      ph.gwt[1] = tb.b2;
      ph.gwt[2] = tb.b3;
    }
    ph.dt = c.dt;
    ph.Ma2 = c.Ma * c.Ma;
    ph.iFr2 = 1.0 / (c.Fr * c.Fr);
    ph.gm = c.gamma_gas;
    ph.planet2 = m.planet2;
  };
  if (stage == 1) {
    PhysPres<1> ph{};
    fill(ph, 1);
    launch(m, ph, dst, nullptr, s);
  } else if (stage == 2) {
    PhysPres<2> ph{};
    fill(ph, 2);
    launch(m, ph, dst, nullptr, s);
  } else {
    PhysPres<3> ph{};
    fill(ph, 3);
    launch(m, ph, dst, nullptr, s);
  }
}

void shell_rhs_velocity(const ShellDev& m, const ShellConsts& c, int stage, const double* const* src, double* dst,
                        hipStream_t s) {
  const long Ns = (long)m.ncell * 8;
  const Tableau tb;
  auto fill = [&](auto& ph, int st) {
    for (int q = 0; q < st; ++q) {
      ph.src[5 * q] = src[3 * q];
      velocity_fields(src[3 * q + 1], Ns, &ph.src[5 * q + 1]);
      ph.src[5 * q + 4] = src[3 * q + 2];
    }
    if (st < 3) ph.src[5 * st] = src[3 * st];
    ph.w = stage_weights(tb, st);
    // gravity: weights of the densities at fields 0, 5 and 10
    if (st == 1) {
      ph.gw[0] = tb.a21t;
      ph.gw[1] = tb.a22t;
      ph.gw[2] = 0.0;
    } else if (st == 2) {
      ph.gw[0] = tb.a31t;
      ph.gw[1] = tb.a32t;
      ph.gw[2] = tb.a33t;
    } else {
      ph.gw[0] = tb.b1;
      ph.gw[1] = tb.b2;
      ph.gw[2] = tb.b3;
    }
    ph.dt = c.dt;
    ph.Ma2 = c.Ma * c.Ma;
    ph.iFr2 = 1.0 / (c.Fr * c.Fr);
    ph.planet2 = m.planet2;
  };
  if (stage == 1) {
    PhysVel<1> ph{};
    fill(ph, 1);
    launch(m, ph, dst, nullptr, s);
  } else if (stage == 2) {
    PhysVel<2> ph{};
    fill(ph, 2);
    launch(m, ph, dst, nullptr, s);
  } else {
    PhysVel<3> ph{};
    fill(ph, 3);
    launch(m, ph, dst, nullptr, s);
  }
}

void shell_B(const ShellDev& m, const ShellConsts& c, double coeff, const double* p, double* dst_u, hipStream_t s) {
  PhysB ph{};
  ph.src[0] = p;
  ph.cf = coeff * c.dt / (c.Ma * c.Ma);
  launch(m, ph, dst_u, nullptr, s);
}

void shell_C(const ShellDev& m, const ShellConsts& c, double coeff, const double* pfix, const double* u,
             double* dst_p, hipStream_t s) {
  const long Ns = (long)m.ncell * 8;
  PhysC<false> ph{};
  ph.src[0] = pfix;
  velocity_fields(u, Ns, &ph.src[1]);
  ph.coeff = coeff;
  ph.dt = c.dt;
  ph.gm = c.gamma_gas;
  launch(m, ph, dst_p, nullptr, s);
}

void shell_d_minus_c(const ShellDev& m, const ShellConsts& c, double coeff, const double* pfix, const double* u,
                     const double* p, const double* scale, double* dst, hipStream_t s) {
  const long Ns = (long)m.ncell * 8;
  PhysC<true> ph{};
  ph.src[0] = p;
  ph.src[1] = pfix;
  velocity_fields(u, Ns, &ph.src[2]);
  ph.coeff = coeff;
  ph.dt = c.dt;
  ph.gm = c.gamma_gas;
  launch(m, ph, dst, scale, s);
}

void shell_b_ainv(const ShellDev& m, const ShellConsts& c, double coeff, const double* p, const double* inv,
                  double* dst_u, hipStream_t s) {
  PhysB ph{};
  ph.src[0] = p;
  ph.cf = coeff * c.dt / (c.Ma * c.Ma);
  ShellEpi ep;
  ep.inv = inv;
  launch<PhysB, 1>(m, ph, dst_u, nullptr, s, ep);
}

void shell_d_minus_c_dots(const ShellDev& m, const ShellConsts& c, double coeff, const double* pfix, const double* u,
                          const double* p, const double* scale, double* dst, const double* V, long ldv, int nv,
                          double* part, long pstride, hipStream_t s) {
  const long Ns = (long)m.ncell * 8;
  const int blocks = (m.ncell + SCPW - 1) / SCPW;
  if (blocks > pstride) throw std::invalid_argument("shell_d_minus_c_dots: partial buffer too small");
  PhysC<true> ph{};
  ph.src[0] = p;
  ph.src[1] = pfix;
  velocity_fields(u, Ns, &ph.src[2]);
  ph.coeff = coeff;
  ph.dt = c.dt;
  ph.gm = c.gamma_gas;
  ShellEpi ep;
  ep.V = V;
  ep.ldv = ldv;
  ep.nv = nv;
  ep.p_last = nv > 0 && p == V + (long)(nv - 1) * ldv;
  ep.part = part;
  ep.pstride = pstride;
  launch<PhysC<true>, 2>(m, ph, dst, scale, s, ep);
}

void shell_mass(const ShellDev& m, int ncomp, const double* rho, double coeff, const double* src, double* dst,
                hipStream_t s) {
  const long Ns = (long)m.ncell * 8;
  if (ncomp == 1 && !rho) {
    PhysMass<1, false> ph{};
    ph.src[0] = src;
    ph.coeff = coeff;
    launch(m, ph, dst, nullptr, s);
  } else if (ncomp == 3 && rho) {
    PhysMass<3, true> ph{};
    velocity_fields(src, Ns, &ph.src[0]);
    ph.src[3] = rho;
    ph.coeff = coeff;
    launch(m, ph, dst, nullptr, s);
  } else if (ncomp == 3) {
    PhysMass<3, false> ph{};
    velocity_fields(src, Ns, &ph.src[0]);
    ph.coeff = coeff;
    launch(m, ph, dst, nullptr, s);
  } else {
    throw std::invalid_argument("shell_mass: 1 component unweighted or 3 components");
  }
}

void shell_block_inverse(const ShellDev& m, const double* rho, double* inv, double* inv_diag, double* bad,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_shell_blocks, dim3((m.ncell + SCPW - 1) / SCPW), dim3(64), 0, s, m, rho, inv, inv_diag, bad);
}

void shell_block_apply(const ShellDev& m, int ncomp, const double* inv, double coeff, const double* b, double* x,
                       hipStream_t s) {
  const long n = (long)m.ncell * 8 * ncomp;
  hipLaunchKernelGGL(k_shell_block_apply, dim3(ceil_div(n, 256)), dim3(256), 0, s, m.ncell, ncomp, inv, coeff, b, x);
}

}  // namespace atmo
