// Spherical-shell mesh of the reference's 3D case (shell.h): coarse panels,
// cell numbering, face neighbours with their orientation, dof positions.
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <stdexcept>

#include "shell.h"

namespace atmo {

namespace {

void unit(double v[3]) {
  const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= n;
  v[1] /= n;
  v[2] /= n;
}

// Corner directions of the coarse panels, [panel][corner s + 2t][3],
// oriented so that d/ds x d/dt points outwards.
std::vector<double> panel_corners(int np) {
  std::vector<double> C;
  auto push = [&](const double* d00, const double* d10, const double* d01, const double* d11) {
    for (const double* d : {d00, d10, d01, d11}) C.insert(C.end(), d, d + 3);
  };
  if (np == 6) {
    // cube face sg * e_a, in-face axes (u, v) with u x v = sg * e_a
    for (int a = 0; a < 3; ++a)
      for (int sg = -1; sg <= 1; sg += 2) {
        int ua = (a + 1) % 3, va = (a + 2) % 3;
        if (sg < 0) std::swap(ua, va);
        double c[4][3];
        for (int t = 0; t < 2; ++t)
          for (int s = 0; s < 2; ++s) {
            double* p = c[s + 2 * t];
            for (int d = 0; d < 3; ++d) p[d] = (d == a ? sg : 0.0) + (d == ua ? 2 * s - 1 : 0) + (d == va ? 2 * t - 1 : 0);
            unit(p);
          }
        push(c[0], c[1], c[2], c[3]);
      }
  } else if (np == 12) {
    // rhombus of the octahedron edge (sa e_a, sb e_b) and the cube corners beside it
    for (int a = 0; a < 3; ++a)
      for (int b = a + 1; b < 3; ++b)
        for (int sa = -1; sa <= 1; sa += 2)
          for (int sb = -1; sb <= 1; sb += 2) {
            const int cx = 3 - a - b;
            double A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, P[3], Q[3];
            A[a] = sa;
            B[b] = sb;
            for (int d = 0; d < 3; ++d) {
              P[d] = A[d] + B[d] + (d == cx ? 1.0 : 0.0);
              Q[d] = A[d] + B[d] - (d == cx ? 1.0 : 0.0);
            }
            unit(P);
            unit(Q);
            // corners A (s,t) = (0,0), P (1,0), Q (0,1), B (1,1); swap P and Q if inward
            const double e1[3] = {P[0] - A[0], P[1] - A[1], P[2] - A[2]};
            const double e2[3] = {Q[0] - A[0], Q[1] - A[1], Q[2] - A[2]};
            const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                  e1[0] * e2[1] - e1[1] * e2[0]};
            if (cr[0] * (A[0] + B[0]) + cr[1] * (A[1] + B[1]) + cr[2] * (A[2] + B[2]) < 0) std::swap(P, Q);
            push(A, P, Q, B);
          }
  } else {
    throw std::invalid_argument("shell: panels must be 6 or 12");
  }
  return C;
}

bool same_dir(const double* a, const double* b) {
  return std::abs(a[0] - b[0]) + std::abs(a[1] - b[1]) + std::abs(a[2] - b[2]) < 1e-12;
}

// the two corners of panel edge f (0: s=0, 1: s=1, 2: t=0, 3: t=1), in the
// direction in which the coordinate along the edge grows
void edge_corners(int f, int& c0, int& c1) {
  static const int E[4][2] = {{0, 2}, {1, 3}, {0, 1}, {2, 3}};
  c0 = E[f][0];
  c1 = E[f][1];
}

// dof at corner (s, t) of face f in the face's coordinates (the two other axes in increasing order)
int face_corner_dof(int f, int s, int t) {
  const int h = f & 1;
  if (f < 2) return h + 2 * s + 4 * t;
  if (f < 4) return s + 2 * h + 4 * t;
  return s + 2 * t + 4 * h;
}

}  // namespace

ShellMesh build_shell(const ShellDesc& d) {
  if (d.refinements < 0 || d.refinements > 7) throw std::invalid_argument("shell: refinements must be in [0, 7]");
  if (!(d.r_inner > 0) || !(d.r_outer > d.r_inner) || !std::isfinite(d.r_outer))
    throw std::invalid_argument("shell: need 0 < r_inner < r_outer");
  if (!(d.planet > 0) || !std::isfinite(d.planet)) throw std::invalid_argument("shell: planet radius must be > 0");
  ShellMesh m;
  m.d = d;
  m.corners = panel_corners(d.panels);
  const int np = d.panels, n = 1 << d.refinements;
  m.n = n;
  m.ncell = np * n * n * n;
  m.nb.assign((size_t)m.ncell * 6, -1);
  m.nbc.assign((size_t)m.ncell * 6, 0);
  // panel edge adjacency: (panel, edge) -> (panel, edge, reversed)
  std::vector<std::array<int, 3>> adj((size_t)np * 4, {-1, -1, 0});
  for (int p = 0; p < np; ++p)
    for (int f = 0; f < 4; ++f) {
      int a0, a1;
      edge_corners(f, a0, a1);
      const double* pa = &m.corners[(size_t)(p * 4 + a0) * 3];
      const double* pb = &m.corners[(size_t)(p * 4 + a1) * 3];
      for (int q = 0; q < np; ++q)
        for (int g = 0; g < 4 && q != p; ++g) {
          int b0, b1;
          edge_corners(g, b0, b1);
          const double* qa = &m.corners[(size_t)(q * 4 + b0) * 3];
          const double* qb = &m.corners[(size_t)(q * 4 + b1) * 3];
          if (same_dir(pa, qa) && same_dir(pb, qb)) adj[p * 4 + f] = {q, g, 0};
          if (same_dir(pa, qb) && same_dir(pb, qa)) adj[p * 4 + f] = {q, g, 1};
        }
      if (adj[p * 4 + f][0] < 0) throw std::logic_error("shell: unmatched panel edge");
    }
  auto id = [&](int p, int i, int j, int k) { return ((p * n + k) * n + j) * n + i; };
  for (int p = 0; p < np; ++p)
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const int c = id(p, i, j, k);
          int* nb = &m.nb[(size_t)c * 6];
          unsigned char* code = &m.nbc[(size_t)c * 6];
          const int ij[4] = {i, i, j, j};  // index across the edge for faces 0..3
          for (int f = 0; f < 4; ++f) {
            const bool inside = (f & 1) ? ij[f] < n - 1 : ij[f] > 0;
            if (inside) {
              const int di = f == 0 ? -1 : (f == 1 ? 1 : 0), dj = f == 2 ? -1 : (f == 3 ? 1 : 0);
              nb[f] = id(p, i + di, j + dj, k);
              code[f] = (unsigned char)(f ^ 1);
              continue;
            }
            const auto& e = adj[p * 4 + f];
            const int q = e[0], g = e[1], rev = e[2];
            const int pos = f < 2 ? j : i;  // position along the edge
            const int pos2 = rev ? n - 1 - pos : pos;
            const int i2 = g < 2 ? (g == 0 ? 0 : n - 1) : pos2;
            const int j2 = g < 2 ? pos2 : (g == 2 ? 0 : n - 1);
            nb[f] = id(q, i2, j2, k);
            code[f] = (unsigned char)(g | (rev << 3));
          }
          if (k > 0) {
            nb[4] = id(p, i, j, k - 1);
            code[4] = 5;
          }
          if (k < n - 1) {
            nb[5] = id(p, i, j, k + 1);
            code[5] = 4;
          }
        }
  // every inner face is seen from both sides with the same orientation, and
  // its corners (s, t) on this side are the corners (flip ? 1 - s : s, t) of
  // the neighbour's face
  std::vector<double> X((size_t)m.ncell * 24);
  shell_dof_points(m, X.data());
  const double tol = 1e-9 * d.r_outer;
  for (int c = 0; c < m.ncell; ++c)
    for (int f = 0; f < 6; ++f) {
      const int o = m.nb[(size_t)c * 6 + f];
      if (o < 0) continue;
      const int g = m.nbc[(size_t)c * 6 + f] & 7, fl = m.nbc[(size_t)c * 6 + f] >> 3;
      if (m.nb[(size_t)o * 6 + g] != c || (m.nbc[(size_t)o * 6 + g] & 7) != f || (m.nbc[(size_t)o * 6 + g] >> 3) != fl)
        throw std::logic_error("shell: face table not symmetric");
      for (int t = 0; t < 2; ++t)
        for (int s = 0; s < 2; ++s) {
          const double* a = &X[((size_t)c * 8 + face_corner_dof(f, s, t)) * 3];
          const double* b = &X[((size_t)o * 8 + face_corner_dof(g, fl ? 1 - s : s, t)) * 3];
          if (std::abs(a[0] - b[0]) + std::abs(a[1] - b[1]) + std::abs(a[2] - b[2]) > tol)
            throw std::logic_error("shell: face orientation does not match the geometry");
        }
    }
  return m;
}

void shell_dof_points(const ShellMesh& m, double* out) {
  const int n = m.n;
  const double dr = (m.d.r_outer - m.d.r_inner) / n;
  for (int c = 0; c < m.ncell; ++c) {
    const int i = c % n, j = (c / n) % n, k = (c / (n * n)) % n, p = c / (n * n * n);
    const double* C = &m.corners[(size_t)p * 12];
    for (int dof = 0; dof < 8; ++dof) {
      const int a = dof & 1, b = (dof >> 1) & 1, cc = dof >> 2;
      const double s = (double)(i + a) / n, t = (double)(j + b) / n;
      double v[3];
      for (int e = 0; e < 3; ++e)
        v[e] = (1 - s) * (1 - t) * C[e] + s * (1 - t) * C[3 + e] + (1 - s) * t * C[6 + e] + s * t * C[9 + e];
      unit(v);
      const double r = m.d.r_inner + dr * (k + cc);
      for (int e = 0; e < 3; ++e) out[((size_t)c * 8 + dof) * 3 + e] = r * v[e];
    }
  }
}

namespace {

double q2v(int a, double x) {
  return a == 0 ? 2.0 * (x - 0.5) * (x - 1.0) : (a == 1 ? -4.0 * x * (x - 1.0) : 2.0 * x * (x - 0.5));
}
double q2d(int a, double x) { return a == 0 ? 4.0 * x - 3.0 : (a == 1 ? 4.0 - 8.0 * x : 4.0 * x - 1.0); }

// x(xi) and det J of the cell map r(zeta) N(xi, eta) (the kernels' map)
double cell_map(const ShellMesh& m, int cell, const double xi[3], double x[3]) {
  const int n = m.n;
  const int i = cell % n, j = (cell / n) % n, k = (cell / (n * n)) % n, p = cell / (n * n * n);
  const double* C = &m.corners[(size_t)p * 12];
  double N[3] = {0, 0, 0}, Nx[3] = {0, 0, 0}, Ny[3] = {0, 0, 0};
  for (int b = 0; b < 3; ++b)
    for (int a = 0; a < 3; ++a) {
      const double s = (i + 0.5 * a) / n, t = (j + 0.5 * b) / n;
      double v[3];
      for (int d = 0; d < 3; ++d)
        v[d] = (1 - s) * (1 - t) * C[d] + s * (1 - t) * C[3 + d] + (1 - s) * t * C[6 + d] + s * t * C[9 + d];
      unit(v);
      const double w = q2v(a, xi[0]) * q2v(b, xi[1]), wx = q2d(a, xi[0]) * q2v(b, xi[1]),
                   wy = q2v(a, xi[0]) * q2d(b, xi[1]);
      for (int d = 0; d < 3; ++d) {
        N[d] += w * v[d];
        Nx[d] += wx * v[d];
        Ny[d] += wy * v[d];
      }
    }
  const double dr = (m.d.r_outer - m.d.r_inner) / n;
  const double r = m.d.r_inner + dr * (k + xi[2]);
  for (int d = 0; d < 3; ++d) x[d] = r * N[d];
  // det [r Nx, r Ny, dr N] = r^2 dr (Nx x Ny) . N
  const double c[3] = {Nx[1] * Ny[2] - Nx[2] * Ny[1], Nx[2] * Ny[0] - Nx[0] * Ny[2], Nx[0] * Ny[1] - Nx[1] * Ny[0]};
  return r * r * dr * (c[0] * N[0] + c[1] * N[1] + c[2] * N[2]);
}

void sbr_at(const SbrParams& sp, const double* x, double& rho, double* u, double& p) {
  const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  const double s2 = x[0] * x[0] + x[1] * x[1];
  const double q = (0.5 * sp.Omega * sp.Omega * s2 - sp.a * sp.g * (1.0 - sp.a / r)) / sp.RT;
  rho = std::exp(q);
  p = std::exp(q);
  u[0] = -sp.Omega * x[1];
  u[1] = sp.Omega * x[0];
  u[2] = 0.0;
}

}  // namespace

void shell_sbr_state(const ShellMesh& m, const SbrParams& sp, double* rho, double* u, double* p) {
  const long Ns = (long)m.ncell * 8;
  std::vector<double> X((size_t)m.ncell * 24);
  shell_dof_points(m, X.data());
  for (long i = 0; i < Ns; ++i) {
    double uu[3];
    sbr_at(sp, &X[(size_t)i * 3], rho[i], uu, p[i]);
    for (int c = 0; c < 3; ++c) u[c * Ns + i] = uu[c];
  }
}

void shell_l2_errors(const ShellMesh& m, const SbrParams& sp, const double* rho, const double* u, const double* p,
                     double* out) {
  const long Ns = (long)m.ncell * 8;
  const double g2[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
  double e[3] = {0, 0, 0}, nrm[3] = {0, 0, 0};
  for (int c = 0; c < m.ncell; ++c)
    for (int q = 0; q < 8; ++q) {
      const double xi[3] = {g2[q & 1], g2[(q >> 1) & 1], g2[q >> 2]};
      double x[3];
      const double JxW = cell_map(m, c, xi, x) * 0.125;
      double re, ue[3], pe;
      sbr_at(sp, x, re, ue, pe);
      double phi[8];
      for (int d = 0; d < 8; ++d)
        phi[d] = ((d & 1) ? xi[0] : 1 - xi[0]) * (((d >> 1) & 1) ? xi[1] : 1 - xi[1]) * ((d >> 2) ? xi[2] : 1 - xi[2]);
      auto val = [&](const double* f) {
        double v = 0;
        for (int d = 0; d < 8; ++d) v += f[(size_t)c * 8 + d] * phi[d];
        return v;
      };
      for (int d = 0; d < 3; ++d) {
        const double v = val(u + d * Ns);
        e[0] += (v - ue[d]) * (v - ue[d]) * JxW;
        nrm[0] += ue[d] * ue[d] * JxW;
      }
      const double rv = val(rho), pv = val(p);
      e[1] += (rv - re) * (rv - re) * JxW;
      nrm[1] += re * re * JxW;
      e[2] += (pv - pe) * (pv - pe) * JxW;
      nrm[2] += pe * pe * JxW;
    }
  for (int i = 0; i < 3; ++i) {
    out[2 * i] = std::sqrt(e[i]);
    out[2 * i + 1] = std::sqrt(e[i]) / std::sqrt(nrm[i]);
  }
}

double shell_min_cell_diameter(const ShellMesh& m) {
  std::vector<double> X((size_t)m.ncell * 24);
  shell_dof_points(m, X.data());
  double best = std::numeric_limits<double>::max();
  for (int c = 0; c < m.ncell; ++c) {
    const double* x = &X[(size_t)c * 24];
    double dmax = 0;
    for (int v = 0; v < 4; ++v) {  // the four diagonals (v, 7 - v)
      const double* a = x + 3 * v;
      const double* b = x + 3 * (7 - v);
      dmax = std::max(dmax, std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                                      (a[2] - b[2]) * (a[2] - b[2])));
    }
    best = std::min(best, dmax);
  }
  return best;
}

}  // namespace atmo
