// Shared definitions of the MI355X (gfx950) IMEX-DG library.
//
// Data layout in HBM: every field is [component][cell][dof] with
// cell = j*nx + i (x fastest) and dof = jj*(k+1) + ii (lexicographic, x
// fastest), so one cell's (k+1)^2 values are contiguous (128 B at Q3 FP64).
//
// Launch shape of the cell kernels ("lane per quadrature column"): one
// wave64 owns CPW = 64 / (2k+1) consecutive cells; lane l works on cell
// l / (2k+1) and on the Gauss column c = l % (2k+1) of that cell, i.e. on
// the 2k+1 quadrature points (c, r), r = 0..2k. The same lane is the face
// quadrature point c of the cell's bottom/top faces and the face
// quadrature point s = c of its left/right faces. Cell dofs are staged
// through LDS with contiguous (coalesced) loads and results leave through
// LDS with contiguous stores.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "basis.h"

namespace atmo {

#define ATMO_HIP_CHECK(expr)                                                                  \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess)                                                                     \
      throw ::atmo::DeviceError(std::string(#expr) + ": " + hipGetErrorString(_e) + " at " + \
                                __FILE__ + ":" + std::to_string(__LINE__));                   \
  } while (0)

struct DeviceError : std::runtime_error {
  explicit DeviceError(const std::string& s) : std::runtime_error(s) {}
};

template <int K>
struct Cfg {
  static constexpr int ND = K + 1;
  static constexpr int ND2 = ND * ND;
  static constexpr int NQ = 2 * K + 1;
  static constexpr int CPW = 64 / NQ;  // cells per wave
  static constexpr int LANES = CPW * NQ;
};

// Terrain-following structured mesh of one MG level: vertex (i,j) at
// x = x0 + i*dx, z = h[i] + (H - h[i]) * j / nz. h has nx+1 entries (for a
// periodic mesh h[nx] == h[0]).
struct MeshDev {
  const double* h;
  int nx, nz;
  double dx, H;
  int periodic;
  int ncell;
};

// Per-cell vertex data (x00,x10,x01,x11 and z likewise).
struct CellVerts {
  double x[4], z[4];
};

__device__ inline CellVerts cell_verts(const MeshDev& m, int i, int j) {
  CellVerts v;
  const double hl = m.h[i], hr = m.h[i + 1];
  const double fj = (double)j / (double)m.nz, fj1 = (double)(j + 1) / (double)m.nz;
  v.x[0] = i * m.dx;
  v.x[1] = (i + 1) * m.dx;
  v.x[2] = v.x[0];
  v.x[3] = v.x[1];
  v.z[0] = hl + (m.H - hl) * fj;
  v.z[1] = hr + (m.H - hr) * fj;
  v.z[2] = hl + (m.H - hl) * fj1;
  v.z[3] = hr + (m.H - hr) * fj1;
  return v;
}

// Bilinear geometry of the lane's quadrature column xi_c. x_eta, z_eta are
// constant along the column; x_xi, z_xi are linear in eta.
template <typename T>
struct ColGeo {
  T xe, ze;      // x_eta(xi_c), z_eta(xi_c)
  T a0, a1;      // x_xi(eta) = a0 + eta*a1
  T b0, b1;      // z_xi(eta) = b0 + eta*b1
  T wc;          // Gauss weight of the column
};

template <int K, typename T>
__device__ inline ColGeo<T> col_geo(const CellVerts& v, int c) {
  constexpr auto& tb = basis::kT<K>;
  double xi = 0, w = 0;
#pragma unroll
  for (int q = 0; q < Cfg<K>::NQ; ++q)
    if (q == c) {
      xi = tb.gq[q];
      w = tb.gw[q];
    }
  ColGeo<T> g;
  g.xe = (T)((1 - xi) * (v.x[2] - v.x[0]) + xi * (v.x[3] - v.x[1]));
  g.ze = (T)((1 - xi) * (v.z[2] - v.z[0]) + xi * (v.z[3] - v.z[1]));
  g.a0 = (T)(v.x[1] - v.x[0]);
  g.a1 = (T)((v.x[3] - v.x[2]) - (v.x[1] - v.x[0]));
  g.b0 = (T)(v.z[1] - v.z[0]);
  g.b1 = (T)((v.z[3] - v.z[2]) - (v.z[1] - v.z[0]));
  g.wc = (T)w;
  return g;
}

// Straight-edge face geometry: outward unit normal and length.
struct FaceGeo {
  double n0, n1, len;
};

// f: 0 left (xi=0), 1 right (xi=1), 2 bottom (eta=0), 3 top (eta=1)
__device__ inline FaceGeo face_geo(const CellVerts& v, int f) {
  double tx, tz, s;
  if (f == 0) { tx = v.x[2] - v.x[0]; tz = v.z[2] - v.z[0]; s = -1; }
  else if (f == 1) { tx = v.x[3] - v.x[1]; tz = v.z[3] - v.z[1]; s = 1; }
  else if (f == 2) { tx = v.x[1] - v.x[0]; tz = v.z[1] - v.z[0]; s = 1; }
  else { tx = v.x[3] - v.x[2]; tz = v.z[3] - v.z[2]; s = -1; }
  FaceGeo g;
  g.len = sqrt(tx * tx + tz * tz);
  // left/top: (-tz, tx)/L ; right/bottom: (tz, -tx)/L
  g.n0 = s * tz / g.len;
  g.n1 = -s * tx / g.len;
  return g;
}

// Lane -> (cell, column) for the lane-per-column layout.
template <int K>
struct LaneMap {
  int lane, cw, c, cell;
  bool active;
  __device__ LaneMap(int ncell) {
    lane = threadIdx.x;
    cw = lane / Cfg<K>::NQ;
    c = lane - cw * Cfg<K>::NQ;
    cell = blockIdx.x * Cfg<K>::CPW + cw;
    active = (cw < Cfg<K>::CPW) && (cell < ncell);
  }
};

// Contiguous staging of the CPW cells of this wave: dst[t] = src[cell0*ND2 + t]
template <int K, typename T, typename S>
__device__ inline void stage_cells(T* __restrict__ lds, const S* __restrict__ src, int ncell) {
  constexpr int N = Cfg<K>::CPW * Cfg<K>::ND2;
  const long base = (long)blockIdx.x * Cfg<K>::CPW * Cfg<K>::ND2;
  const long lim = (long)ncell * Cfg<K>::ND2;
  for (int t = threadIdx.x; t < N; t += 64) {
    long g = base + t;
    lds[t] = (g < lim) ? (T)src[g] : (T)0;
  }
}

// Per-lane copies of the lane-dependent table rows (select chains, no memory).
template <int K, typename T>
__device__ inline void lane_rows(int c, T (&Sc)[Cfg<K>::ND], T (&Dc)[Cfg<K>::ND]) {
  constexpr auto& tb = basis::kT<K>;
#pragma unroll
  for (int i = 0; i < Cfg<K>::ND; ++i) {
    Sc[i] = 0;
    Dc[i] = 0;
  }
#pragma unroll
  for (int q = 0; q < Cfg<K>::NQ; ++q)
    if (q == c) {
#pragma unroll
      for (int i = 0; i < Cfg<K>::ND; ++i) {
        Sc[i] = (T)tb.S[q][i];
        Dc[i] = (T)tb.D[q][i];
      }
    }
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace atmo
