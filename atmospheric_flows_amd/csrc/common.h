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

// Mesh of one MG level. Structured terrain-following form: vertex (i,j) at
// x = i*dx, z = h[i] + (H - h[i]) * j * inv_nz, cell = j*nx + i; h has nx+1
// entries (for a periodic mesh h[nx] == h[0]). Locally refined form (cgeo
// set): per-cell geometry cgeo[6*cell] = (x0, dx, z00, z10, z01, z11) of
// the cell's bilinear map and, on the finest level, the face table cface.
// Either way cell columns are vertical: every cell's left and right edges
// have constant x, so x_eta = 0 and x_xi = dx exactly.
//
// cface[16*cell + 4*f + {0,1,2,3}] = (type, n0, n1, half) for face f (0 left,
// 1 right, 2 bottom, 3 top): FACE_BDRY; FACE_SAME (neighbour n0 of the same
// size); FACE_COARSER (this face is half `half` of face f^1 of the larger
// neighbour n0); FACE_FINER (the face is split: half 0 is a face of n0,
// half 1 of n1).
//
// Transfers to the next finer level: child[4*cell + a + 2b] (child (a,b) of
// a refined cell, or (same cell, -1, -1, -1)); from the next coarser level:
// parent[2*cell] = (coarse cell, a + 2b or -1 for the same cell). Null:
// the structured rule (children 2I+a, 2J+b).
enum FaceType { FACE_BDRY = 0, FACE_SAME = 1, FACE_COARSER = 2, FACE_FINER = 3 };

struct MeshDev {
  const double* h;
  int nx, nz;
  double dx, H, inv_nz;
  int periodic;
  int ncell;
  const double* cgeo;
  const int* cface;
  const int* child;
  const int* parent;
};

// The 1D tables in constant memory, for lane-dependent rows and columns.
template <int K>
__constant__ basis::Tables<K> cT = basis::make_tables<K>();

// Per-cell vertex data (x00,x10,x01,x11 and z likewise) and the cell width.
struct CellVerts {
  double x[4], z[4];
  double dx;
};

// The loaded words a cell's vertices are formed from: the six of cgeo on a
// locally refined mesh, else the heights h[i], h[i+1] of its column. Kernels
// load them before their other loads (cell_verts_load) and form the vertices
// after issuing those (cell_verts_form), so the loads share one memory
// latency instead of waiting for the heights first.
struct VertsRaw {
  double g[6];
};

__device__ inline VertsRaw cell_verts_load(const MeshDev& m, int cell) {
  VertsRaw r;
  if (m.cgeo) {
    const double* g = m.cgeo + 6 * (long)cell;
#pragma unroll
    for (int q = 0; q < 6; ++q) r.g[q] = g[q];
  } else {
    const int i = cell % m.nx;
    r.g[0] = m.h[i];
    r.g[1] = m.h[i + 1];
#pragma unroll
    for (int q = 2; q < 6; ++q) r.g[q] = 0.0;
  }
  return r;
}

__device__ inline CellVerts cell_verts_form(const MeshDev& m, int cell, const VertsRaw& r) {
  CellVerts v;
  if (m.cgeo) {
    const double* g = r.g;
    v.dx = g[1];
    v.x[0] = v.x[2] = g[0];
    v.x[1] = v.x[3] = g[0] + g[1];
    v.z[0] = g[2];
    v.z[1] = g[3];
    v.z[2] = g[4];
    v.z[3] = g[5];
    return v;
  }
  const int j = cell / m.nx, i = cell - j * m.nx;
  v.dx = m.dx;
  const double hl = r.g[0], hr = r.g[1];
  const double fj = j * m.inv_nz, fj1 = (j + 1) * m.inv_nz;
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

__device__ inline CellVerts cell_verts(const MeshDev& m, int cell) {
  return cell_verts_form(m, cell, cell_verts_load(m, cell));
}

// Bilinear geometry of the lane's quadrature column xi_c on a vertical cell
// column: x_xi = dx, x_eta = 0, z_eta(xi_c) = ze (constant along the
// column) and z_xi(eta) = b0 + eta*b1. Then det J = dx*ze, and for a
// physical gradient (Gx, Gz) the reference-space integrands are
// G_xi = ze*Gx and G_eta = dx*Gz - z_xi*Gx (cofactor form).
template <typename T>
struct ColGeo {
  T dx, ze, b0, b1;
  T wc;  // Gauss weight of the column
};

template <int K, typename T>
__device__ inline ColGeo<T> col_geo(const CellVerts& v, int c) {
  const double xi = cT<K>.gq[c];
  ColGeo<T> g;
  g.dx = (T)v.dx;
  g.ze = (T)((1 - xi) * (v.z[2] - v.z[0]) + xi * (v.z[3] - v.z[1]));
  g.b0 = (T)(v.z[1] - v.z[0]);
  g.b1 = (T)((v.z[3] - v.z[2]) - (v.z[1] - v.z[0]));
  g.wc = (T)cT<K>.gw[c];
  return g;
}

// Straight-edge face geometry: outward unit normal and length.
struct FaceGeo {
  double n0, n1, len;
};

// f: 0 left (xi=0), 1 right (xi=1), 2 bottom (eta=0), 3 top (eta=1). The
// left/right edges are vertical, so their normals are exactly (-1,0)/(1,0).
__device__ inline FaceGeo face_geo(const CellVerts& v, int f) {
  const double dx = v.dx;
  FaceGeo g;
  if (f < 2) {
    g.len = f == 0 ? v.z[2] - v.z[0] : v.z[3] - v.z[1];
    g.n0 = f == 0 ? -1.0 : 1.0;
    g.n1 = 0.0;
  } else {
    const double tz = f == 2 ? v.z[1] - v.z[0] : v.z[3] - v.z[2];
    g.len = sqrt(dx * dx + tz * tz);
    const double s = (f == 2 ? 1.0 : -1.0) / g.len;  // bottom: (tz, -dx)/L ; top: (-tz, dx)/L
    g.n0 = s * tz;
    g.n1 = -s * dx;
  }
  return g;
}

// Lane -> (cell, column) for the lane-per-column layout.
// Block of cells a one-wave workgroup of a cell kernel handles. Workgroups
// are dispatched round-robin over the chip's 8 XCDs, each with its own L2;
// this numbering gives each XCD one contiguous range of blocks, so the
// cells above and below a cell (one mesh row away) are mostly handled by
// the same XCD shortly before or after it and their face traces are L2 hits.
// A bijection of [0, gridDim.x).
__device__ inline int cell_block() {
  const int G = gridDim.x, b = blockIdx.x, x = b & 7, r = b >> 3;
  const int q = G >> 3, rem = G & 7;
  return x * q + (x < rem ? x : rem) + r;
}

template <int K>
struct LaneMap {
  int lane, cw, c, cell;
  bool active;
  __device__ LaneMap(int ncell) {
    lane = threadIdx.x;
    cw = lane / Cfg<K>::NQ;
    c = lane - cw * Cfg<K>::NQ;
    cell = cell_block() * Cfg<K>::CPW + cw;
    active = (cw < Cfg<K>::CPW) && (cell < ncell);
  }
};

// Staging of NF fields of the CPW cells of this wave (64 threads): dof d of
// block cell w of field f goes to lds[f*CPW*STRIDE + w*STRIDE + d] (STRIDE >
// ND2 pads the cells apart in the LDS banks). The passes have a fixed trip
// count and every load of a group of up to four fields is issued before the
// group's first LDS write, so the group costs one memory latency.
// HALO = 1: the cells just before and after the block's are staged too (at
// w = 0 and CPW + 1, the block's cells at w + 1; cells outside the mesh read
// as zeros).
template <int K, int NF, typename T, typename S, int STRIDE = Cfg<K>::ND2, int HALO = 0>
__device__ inline void stage_fields(T* __restrict__ lds, const S* const (&src)[NF], int ncell) {
  constexpr int ND2 = Cfg<K>::ND2, NC = Cfg<K>::CPW + 2 * HALO, N = NC * ND2, NPASS = (N + 63) / 64;
  constexpr int FS = NC * STRIDE;
  constexpr int G = 4;
  const long base = (long)cell_block() * Cfg<K>::CPW * ND2 - HALO * ND2;
  const long lim = (long)ncell * ND2;
#pragma unroll
  for (int f0 = 0; f0 < NF; f0 += G) {
    T v[G][NPASS];
#pragma unroll
    for (int f = 0; f < G; ++f)
#pragma unroll
      for (int it = 0; it < NPASS; ++it) {
        const int t = threadIdx.x + it * 64;
        const long g = base + t;
        v[f][it] = (f0 + f < NF && t < N && g >= 0 && g < lim) ? (T)src[f0 + f < NF ? f0 + f : 0][g] : (T)0;
      }
#pragma unroll
    for (int f = 0; f < G; ++f)
#pragma unroll
      for (int it = 0; it < NPASS; ++it) {
        const int t = threadIdx.x + it * 64;
        if (f0 + f < NF && t < N) lds[(f0 + f) * FS + (STRIDE == ND2 ? t : (t / ND2) * STRIDE + t % ND2)] = v[f][it];
      }
  }
}

// LDS stride of one staged cell: ND2 rounded up to an odd number of
// doubles, so the CPW cells of a wave start in different banks.
template <int K>
constexpr int lds_cell_stride() {
  return Cfg<K>::ND2 % 2 == 0 ? Cfg<K>::ND2 + 1 : Cfg<K>::ND2;
}

// Symmetric ND2 x ND2 cell blocks stored as their upper triangle, row by
// row: SIZE doubles per cell, entry (a, b) = (b, a) at index(a, b).
// LDS_STRIDE (odd) spaces staged blocks apart in the LDS banks.
template <int K>
struct SymBlock {
  static constexpr int N = Cfg<K>::ND2, SIZE = N * (N + 1) / 2, LDS_STRIDE = SIZE | 1;
  __host__ __device__ static constexpr int index(int a, int b) {
    return a <= b ? a * (2 * N - a + 1) / 2 + (b - a) : b * (2 * N - b + 1) / 2 + (a - b);
  }
};

// Per-lane copies of the lane-dependent table rows Sc[i] = S[c][i],
// Dc[i] = D[c][i] (constant-memory loads).
template <int K, typename T>
__device__ inline void lane_rows(int c, T (&Sc)[Cfg<K>::ND], T (&Dc)[Cfg<K>::ND]) {
#pragma unroll
  for (int i = 0; i < Cfg<K>::ND; ++i) {
    Sc[i] = (T)cT<K>.S[c][i];
    Dc[i] = (T)cT<K>.D[c][i];
  }
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace atmo
