// Compile-time 1D tables of the DG discretization:
//  - FE_DGQ(k) support points = Gauss-Lobatto nodes on [0,1]
//    (FE_DGQ as used at Atmospheric_Flow.cc:227-229),
//  - QGauss<1>(2k+1) points and weights on [0,1] (Atmospheric_Flow.cc:345),
//  - Lagrange basis values S[q][i] = l_i(x_q) and derivatives D[q][i].
// Everything is evaluated by constexpr Newton iterations, so kernels see the
// tables as literals after unrolling.
#pragma once

namespace atmo {
namespace basis {

constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double cabs(double x) { return x < 0 ? -x : x; }

// cos on [0, pi] by Taylor series (argument range used here only).
constexpr double ccos(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 40; ++n) {
    term *= -x2 / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}

// P_n(z) and P_{n-1}(z) by the three-term recurrence
constexpr void legendre(int n, double z, double& p, double& pm1) {
  double p0 = 1.0, p1 = z;
  if (n == 0) {
    p = 1.0;
    pm1 = 0.0;
    return;
  }
  for (int m = 2; m <= n; ++m) {
    double p2 = ((2.0 * m - 1.0) * z * p1 - (m - 1.0) * p0) / m;
    p0 = p1;
    p1 = p2;
  }
  p = p1;
  pm1 = p0;
}

template <int K>
struct Tables {
  static constexpr int ND = K + 1;
  static constexpr int NQ = 2 * K + 1;
  double gll[ND];
  double gq[NQ];
  double gw[NQ];
  double S[NQ][ND];
  double D[NQ][ND];
  double St[ND][NQ];  // transposes: St[i][q] = S[q][i]
  double Dt[ND][NQ];
  double Sh[2][NQ][ND];   // Sh[h][q][i] = l_i((h + gq[q])/2): points of half h of a split face
  double ShT[2][ND][NQ];  // its transpose
};

template <int K>
constexpr Tables<K> make_tables() {
  Tables<K> t{};
  constexpr int ND = K + 1, NQ = 2 * K + 1;
  // Gauss-Legendre
  for (int i = 0; i < NQ; ++i) {
    double z = ccos(kPi * (i + 0.75) / (NQ + 0.5));
    double p = 0, pm1 = 0, dp = 1;
    for (int it = 0; it < 100; ++it) {
      legendre(NQ, z, p, pm1);
      dp = NQ * (z * p - pm1) / (z * z - 1.0);
      double dz = p / dp;
      z -= dz;
      if (cabs(dz) < 1e-17) break;
    }
    legendre(NQ, z, p, pm1);
    dp = NQ * (z * p - pm1) / (z * z - 1.0);
    t.gq[i] = (1.0 - z) / 2.0;
    t.gw[i] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
  // Gauss-Lobatto: endpoints and roots of P_K'
  t.gll[0] = 0.0;
  t.gll[K] = 1.0;
  for (int i = 1; i < K; ++i) {
    double z = ccos(kPi * i / K);
    for (int it = 0; it < 100; ++it) {
      double p = 0, pm1 = 0;
      legendre(K, z, p, pm1);
      double dp = K * (z * p - pm1) / (z * z - 1.0);
      double d2p = (2.0 * z * dp - K * (K + 1.0) * p) / (1.0 - z * z);
      double dz = dp / d2p;
      z -= dz;
      if (cabs(dz) < 1e-17) break;
    }
    t.gll[i] = (1.0 - z) / 2.0;
  }
  // Lagrange values and derivatives at the Gauss points
  for (int q = 0; q < NQ; ++q) {
    const double x = t.gq[q];
    for (int i = 0; i < ND; ++i) {
      double v = 1.0;
      for (int m = 0; m < ND; ++m)
        if (m != i) v *= (x - t.gll[m]) / (t.gll[i] - t.gll[m]);
      t.S[q][i] = v;
      double s = 0.0;
      for (int l = 0; l < ND; ++l) {
        if (l == i) continue;
        double pr = 1.0 / (t.gll[i] - t.gll[l]);
        for (int m = 0; m < ND; ++m)
          if (m != i && m != l) pr *= (x - t.gll[m]) / (t.gll[i] - t.gll[m]);
        s += pr;
      }
      t.D[q][i] = s;
      t.St[i][q] = t.S[q][i];
      t.Dt[i][q] = s;
      for (int h = 0; h < 2; ++h) {
        const double xh = 0.5 * (h + x);
        double vh = 1.0;
        for (int m = 0; m < ND; ++m)
          if (m != i) vh *= (xh - t.gll[m]) / (t.gll[i] - t.gll[m]);
        t.Sh[h][q][i] = vh;
        t.ShT[h][i][q] = vh;
      }
    }
  }
  return t;
}

// 1D embedding of a parent cell into its children (DG prolongation):
// P[a][i][n] = l_n((a + x_i)/2), a = child half, i = child node, n = parent node.
// R[a][m][i] = 1/2 (M^{-1} P_a^T M)[m][i], the L2 projection child -> parent.
template <int K>
struct Transfer {
  static constexpr int ND = K + 1;
  double P[2][ND][ND];
  double R[2][ND][ND];
};

template <int K>
constexpr Transfer<K> make_transfer() {
  constexpr int ND = K + 1, NQ = 2 * K + 1;
  constexpr Tables<K> t = make_tables<K>();
  Transfer<K> tr{};
  double M[ND][ND] = {};
  for (int m = 0; m < ND; ++m)
    for (int n = 0; n < ND; ++n) {
      double s = 0;
      for (int q = 0; q < NQ; ++q) s += t.gw[q] * t.S[q][m] * t.S[q][n];
      M[m][n] = s;
    }
  // inverse of M (SPD, Gauss-Jordan without pivoting is stable here)
  double A[ND][ND] = {}, Mi[ND][ND] = {};
  for (int i = 0; i < ND; ++i)
    for (int j = 0; j < ND; ++j) {
      A[i][j] = M[i][j];
      Mi[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < ND; ++c) {
    double d = A[c][c];
    for (int l = 0; l < ND; ++l) {
      A[c][l] /= d;
      Mi[c][l] /= d;
    }
    for (int r = 0; r < ND; ++r)
      if (r != c) {
        double f = A[r][c];
        for (int l = 0; l < ND; ++l) {
          A[r][l] -= f * A[c][l];
          Mi[r][l] -= f * Mi[c][l];
        }
      }
  }
  for (int a = 0; a < 2; ++a) {
    for (int i = 0; i < ND; ++i)
      for (int n = 0; n < ND; ++n) {
        const double x = 0.5 * (a + t.gll[i]);
        double v = 1.0;
        for (int m = 0; m < ND; ++m)
          if (m != n) v *= (x - t.gll[m]) / (t.gll[n] - t.gll[m]);
        tr.P[a][i][n] = v;
      }
    double PtM[ND][ND] = {};
    for (int n = 0; n < ND; ++n)
      for (int i = 0; i < ND; ++i) {
        double s = 0;
        for (int l = 0; l < ND; ++l) s += tr.P[a][l][n] * M[l][i];
        PtM[n][i] = s;
      }
    for (int m = 0; m < ND; ++m)
      for (int i = 0; i < ND; ++i) {
        double s = 0;
        for (int n = 0; n < ND; ++n) s += Mi[m][n] * PtM[n][i];
        tr.R[a][m][i] = 0.5 * s;
      }
  }
  return tr;
}

template <int K>
inline constexpr Tables<K> kT = make_tables<K>();
template <int K>
inline constexpr Transfer<K> kTr = make_transfer<K>();

}  // namespace basis
}  // namespace atmo
