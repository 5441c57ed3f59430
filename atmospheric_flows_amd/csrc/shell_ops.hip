// Operator kernels of the spherical-shell path (shell.h): the EULEROperator
// forms of euler_operator.h in dim = 3 at FE_DGQ(1) with QGauss(3)
// (Atmospheric_Flow.cc:345), on MappingQ(2)-mapped cells, radial gravity
// g a^2/r^2 e_r evaluated at every quadrature point.
//
// Launch shape: one wave64 per two cells; lanes 0-31 work on the first,
// 32-63 on the second. Lane l < 27 of a cell owns quadrature point
// (l % 3, l / 3 % 3, l / 9) in the cell phase and face points l and l + 27
// (face (l + 27h) / 9, point (l + 27h) % 9) in the face phase; lanes
// l < 8 * NOUT then integrate test function l % 8 of output component l / 8
// from the LDS records of both phases. The cell's dofs are staged through
// LDS; a neighbour's dofs are read from global memory (L2) at the face
// point its orientation code maps to. J is rebuilt from the cell's 9
// lateral support directions (formed once per cell by lanes 0-8) and its
// layer radii, so no per-point geometry is stored.
#include <cmath>

#include "ops.h"
#include "shell.h"

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

// x(xi) and J[d][e] = dx_d / dxi_e of the cell map r(zeta) N(xi, eta)
__device__ inline void shell_map(const double (*N)[3], double r0, double dr, const double xi[3], double x[3],
                                 double J[3][3]) {
  double La[3], Da[3], Lb[3], Db[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    La[a] = q2v(a, xi[0]);
    Da[a] = q2d(a, xi[0]);
    Lb[a] = q2v(a, xi[1]);
    Db[a] = q2d(a, xi[1]);
  }
  double n[3] = {0, 0, 0}, nx[3] = {0, 0, 0}, ny[3] = {0, 0, 0};
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double w = La[a] * Lb[b], wx = Da[a] * Lb[b], wy = La[a] * Db[b];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double v = N[a + 3 * b][d];
        n[d] += w * v;
        nx[d] += wx * v;
        ny[d] += wy * v;
      }
    }
  const double r = r0 + dr * xi[2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x[d] = r * n[d];
    J[d][0] = r * nx[d];
    J[d][1] = r * ny[d];
    J[d][2] = dr * n[d];
  }
}

// det J and Ji[e][d] = dxi_e / dx_d
__device__ inline double inv3(const double J[3][3], double Ji[3][3]) {
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / det;
  Ji[0][0] = c00 * id;
  Ji[1][0] = c01 * id;
  Ji[2][0] = c02 * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  return det;
}

// trilinear basis values at xi
__device__ inline void q1_all(const double xi[3], double phi[8]) {
#pragma unroll
  for (int d = 0; d < 8; ++d) phi[d] = q1v(d & 1, xi[0]) * q1v((d >> 1) & 1, xi[1]) * q1v(d >> 2, xi[2]);
}

// reference point of face f at face Gauss point (gs, gt)
__device__ inline void face_xi(int f, int gs, int gt, double xi[3]) {
  const int e = f >> 1;
  xi[e] = (double)(f & 1);
  xi[e == 0 ? 1 : 0] = kGQ[gs];
  xi[e == 2 ? 1 : 2] = kGQ[gt];
}

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
  double gwt[3];
  int gsl[3];  // gravity work terms: weight and triple (-1: the kinetic pair)
  int ngr;
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
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      if (t >= ngr) break;
      const int sl = gsl[t];
      const double rr = sl < 0 ? rk : v[5 * sl];
      const double ue = dot3(sl < 0 ? uk : v + 5 * sl + 1, er);
      val -= gwt[t] * dt * Ma2 * iFr2 * rr * ue * G;
    }
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
// Lateral support directions of the cell (lanes 0-8) into N, its inner radius into r0.
__device__ inline void cell_directions(const ShellDev& m, int cell, int l, double (*N)[3], double* r0) {
  const int n = m.n;
  const int i = cell % n, j = (cell / n) % n, k = (cell / (n * n)) % n, p = cell / (n * n * n);
  if (l < 9) {
    const int a = l % 3, b = l / 3;
    const double s = (i + 0.5 * a) / n, t = (j + 0.5 * b) / n;
    const double* C = m.corners + 12 * p;
    double v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      v[d] = (1 - s) * (1 - t) * C[d] + s * (1 - t) * C[3 + d] + (1 - s) * t * C[6 + d] + s * t * C[9 + d];
    const double nn = sqrt(dot3(v, v));
#pragma unroll
    for (int d = 0; d < 3; ++d) N[l][d] = v[d] / nn;
  }
  if (l == 0) *r0 = m.r_in + m.dr * k;
}

template <class Phys>
__global__ void __launch_bounds__(64) k_shell(ShellDev m, Phys ph, double* __restrict__ dst, long Ns,
                                              const double* __restrict__ scale) {
  constexpr int NF = Phys::NF, NO = Phys::NOUT;
  __shared__ double sdof[2][NF][8];
  __shared__ double sN[2][9][3];
  __shared__ double sr0[2];
  __shared__ double svol[2][NO][27][4];
  __shared__ double sface[2][NO][Phys::FACES ? 54 : 1];
  const int lane = threadIdx.x, w = lane >> 5, l = lane & 31;
  const int c0 = cell_block() * 2;
  const int cell = c0 + w;
  const bool active = cell < m.ncell;
  // stage the dofs of both cells (16 contiguous doubles per field)
  const long lim = (long)m.ncell * 8;
  for (int t = lane; t < 16 * NF; t += 64) {
    const int f = t >> 4, r = t & 15;
    const long g = (long)c0 * 8 + r;
    sdof[r >> 3][f][r & 7] = g < lim ? ph.src[f][g] : 0.0;
  }
  if (active) cell_directions(m, cell, l, sN[w], &sr0[w]);
  __syncthreads();
  if (active && l < 27) {
    const double xi[3] = {kGQ[l % 3], kGQ[(l / 3) % 3], kGQ[l / 9]};
    double x[3], J[3][3], Ji[3][3], phi[8], v[NF];
    shell_map(sN[w], sr0[w], m.dr, xi, x, J);
    const double JxW = inv3(J, Ji) * kGW[l % 3] * kGW[(l / 3) % 3] * kGW[l / 9];
    q1_all(xi, phi);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double a = 0;
#pragma unroll
      for (int d = 0; d < 8; ++d) a += sdof[w][f][d] * phi[d];
      v[f] = a;
    }
    double V[NO], F[NO][3];
    ph.cell(v, x, V, F);
#pragma unroll
    for (int o = 0; o < NO; ++o) {
      svol[w][o][l][0] = V[o] * JxW;
#pragma unroll
      for (int e = 0; e < 3; ++e) svol[w][o][l][1 + e] = (Ji[e][0] * F[o][0] + Ji[e][1] * F[o][1] + Ji[e][2] * F[o][2]) * JxW;
    }
  }
  if (Phys::FACES && active && l < 27) {
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const int pt = l + 27 * h, f = pt / 9, s = pt % 9, gs = s % 3, gt = s / 3, e = f >> 1;
      double xi[3], x[3], J[3][3], Ji[3][3], phi[8];
      face_xi(f, gs, gt, xi);
      shell_map(sN[w], sr0[w], m.dr, xi, x, J);
      const double det = inv3(J, Ji);
      const double sg = (f & 1) ? det : -det;
      double nv[3] = {sg * Ji[e][0], sg * Ji[e][1], sg * Ji[e][2]};
      const double area = sqrt(dot3(nv, nv));
      const double nrm[3] = {nv[0] / area, nv[1] / area, nv[2] / area};
      const double JxW = area * kGW[gs] * kGW[gt];
      q1_all(xi, phi);
      double vp[NF], vm[NF], Fv[NO];
#pragma unroll
      for (int q = 0; q < NF; ++q) {
        double a = 0;
#pragma unroll
        for (int d = 0; d < 8; ++d) a += sdof[w][q][d] * phi[d];
        vp[q] = a;
      }
      const int nbr = m.nb[(long)cell * 6 + f];
      if (nbr < 0) {
        ph.wall(vp, nrm, Fv);
      } else {
        const int code = m.nbc[(long)cell * 6 + f];
        const int g = code & 7, gs2 = (code >> 3) ? 2 - gs : gs;
        double xn[3], pn[8];
        face_xi(g, gs2, gt, xn);
        q1_all(xn, pn);
        // the neighbour's dofs on its face g: index along axis g/2 equal to g & 1
        const int ax = g >> 1, side = g & 1;
#pragma unroll
        for (int q = 0; q < NF; ++q) {
          const double* src = ph.src[q] + (long)nbr * 8;
          double a = 0;
#pragma unroll
          for (int d = 0; d < 8; ++d)
            if (((d >> ax) & 1) == side) a += src[d] * pn[d];
          vm[q] = a;
        }
        ph.face(vp, vm, nrm, Fv);
      }
#pragma unroll
      for (int o = 0; o < NO; ++o) sface[w][o][Phys::FACES ? pt : 0] = Fv[o] * JxW;
    }
  }
  __syncthreads();
  if (active && l < 8 * NO) {
    const int i = l & 7, o = l >> 3;
    const int ia = i & 1, ib = (i >> 1) & 1, ic = i >> 2;
    double acc = 0;
#pragma unroll 3
    for (int q = 0; q < 27; ++q) {
      const int q0 = q % 3, q1 = (q / 3) % 3, q2 = q / 9;
      const double la = q1v(ia, kGQ[q0]), lb = q1v(ib, kGQ[q1]), lc = q1v(ic, kGQ[q2]);
      acc += svol[w][o][q][0] * la * lb * lc + svol[w][o][q][1] * q1d(ia) * lb * lc +
             svol[w][o][q][2] * la * q1d(ib) * lc + svol[w][o][q][3] * la * lb * q1d(ic);
    }
    if (Phys::FACES) {
      const int idx[3] = {ia, ib, ic};
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const int e = f >> 1;
        if (idx[e] != (f & 1)) continue;  // phi_i vanishes on this face
        const int s0 = idx[e == 0 ? 1 : 0], s1 = idx[e == 2 ? 1 : 2];
#pragma unroll
        for (int s = 0; s < 9; ++s)
          acc += sface[w][o][Phys::FACES ? f * 9 + s : 0] * q1v(s0, kGQ[s % 3]) * q1v(s1, kGQ[s / 3]);
      }
    }
    const long k = (long)o * Ns + (long)cell * 8 + i;
    dst[k] = scale ? acc * scale[k] : acc;
  }
}

template <class Phys>
void launch(const ShellDev& m, const Phys& ph, double* dst, const double* scale, hipStream_t s) {
  const int blocks = (m.ncell + 1) / 2;
  hipLaunchKernelGGL(k_shell<Phys>, dim3(blocks), dim3(64), 0, s, m, ph, dst, (long)m.ncell * 8, scale);
}

// ------------------------------------------------------------ cell blocks
// Packed upper triangle of a symmetric 8 x 8 block: entry (a, b), a <= b.
__host__ __device__ constexpr int sym8(int a, int b) { return a <= b ? a * (15 - a) / 2 + b : b * (15 - b) / 2 + a; }

// Inverse of each cell's mass block (weighted by rho when given), by
// Cholesky in one lane per cell; the inverse diagonal of the block itself too.
__global__ void __launch_bounds__(64) k_shell_blocks(ShellDev m, const double* __restrict__ rho,
                                                     double* __restrict__ inv, double* __restrict__ inv_diag,
                                                     double* __restrict__ bad) {
  __shared__ double sN[2][9][3];
  __shared__ double sr0[2];
  __shared__ double swq[2][27];
  __shared__ double sB[2][36];
  const int lane = threadIdx.x, w = lane >> 5, l = lane & 31;
  const int cell = cell_block() * 2 + w;
  const bool active = cell < m.ncell;
  if (active) cell_directions(m, cell, l, sN[w], &sr0[w]);
  __syncthreads();
  if (active && l < 27) {
    const double xi[3] = {kGQ[l % 3], kGQ[(l / 3) % 3], kGQ[l / 9]};
    double x[3], J[3][3], Ji[3][3], phi[8];
    shell_map(sN[w], sr0[w], m.dr, xi, x, J);
    double wq = inv3(J, Ji) * kGW[l % 3] * kGW[(l / 3) % 3] * kGW[l / 9];
    if (rho) {
      q1_all(xi, phi);
      double r = 0;
#pragma unroll
      for (int d = 0; d < 8; ++d) r += rho[(long)cell * 8 + d] * phi[d];
      wq *= r;
    }
    swq[w][l] = wq;
  }
  __syncthreads();
  if (active) {
    for (int e = l; e < 36; e += 32) {
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
  if (active && l == 0) {
    double L[8][8];
    bool spd = true;
    for (int j = 0; j < 8; ++j) {
      double s = sB[w][sym8(j, j)];
      for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
      spd = spd && s > 0.0;  // fails for a density that is not positive in the cell (or NaN)
      L[j][j] = sqrt(s);
      for (int i = j + 1; i < 8; ++i) {
        double t = sB[w][sym8(j, i)];
        for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
        L[i][j] = t / L[j][j];
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
        double t = i == c ? 1.0 : 0.0;
        for (int k = c; k < i; ++k) t -= L[i][k] * Li[k][c];
        Li[i][c] = t / L[i][i];
      }
    for (int a = 0; a < 8; ++a)
      for (int b = a; b < 8; ++b) {
        double t = 0;
        for (int k = b; k < 8; ++k) t += Li[k][a] * Li[k][b];
        inv[(long)cell * 36 + sym8(a, b)] = t;
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
    if (st == 1) {
      ph.gwt[0] = tb.a21t;
      ph.gsl[0] = 0;
      ph.gwt[1] = tb.a22t;
      ph.gsl[1] = -1;
      ph.ngr = 2;
    } else if (st == 2) {
      ph.gwt[0] = tb.a31t;
      ph.gsl[0] = 0;
      ph.gwt[1] = tb.a32t;
      ph.gsl[1] = 1;
      ph.gwt[2] = tb.a33t;
      ph.gsl[2] = -1;
      ph.ngr = 3;
    } else {
      ph.gwt[0] = tb.b1;
      ph.gsl[0] = 0;
      ph.gwt[1] = tb.b2;
      ph.gsl[1] = 1;
      ph.gwt[2] = tb.b3;
      ph.gsl[2] = 2;
      ph.ngr = 3;
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
  hipLaunchKernelGGL(k_shell_blocks, dim3((m.ncell + 1) / 2), dim3(64), 0, s, m, rho, inv, inv_diag, bad);
}

void shell_block_apply(const ShellDev& m, int ncomp, const double* inv, double coeff, const double* b, double* x,
                       hipStream_t s) {
  const long n = (long)m.ncell * 8 * ncomp;
  hipLaunchKernelGGL(k_shell_block_apply, dim3(ceil_div(n, 256)), dim3(256), 0, s, m.ncell, ncomp, inv, coeff, b, x);
}

}  // namespace atmo
