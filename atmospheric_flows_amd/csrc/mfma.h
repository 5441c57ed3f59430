// Mass-type cell operators on the matrix cores (FP32 MFMA), for the FP32
// multigrid levels.
//
// Layout ("4 lanes per cell"): a wave64 holds 16 cells; lane l works on cell
// b = l / 4 of the wave and on dof row j = l % 4 of that cell, and keeps the
// row's dofs (j, i), i = 0..3, in four registers (rows and columns past k
// are zero padding, so k = 2 and k = 3 share the code).
//
// v_mfma_f32_4x4x1_16b_f32 computes 16 independent 4x4 outer products, one
// per group of four lanes, i.e. one per cell: D[m][n] += A[m] B[n] with A[m]
// from lane 4b+m, B[n] from lane 4b+n, and D[m][n] in register m of lane
// 4b+n. The four 1D contractions of the sum-factorized mass apply
// M u = (S x S)^T diag(w) (S x S) u are chained so that every result is
// already in the operand layout of the next one (no shuffles, no LDS):
//   1. X1[j][qx]  = sum_i u[j][i] S[qx][i]     A = u (register i),   B = S rows
//   2. X2[qz][qx] = sum_j S[qz][j] X1[j][qx]   A = S rows,           B = X1 (register j)
//   3. X2 *= w(qz, qx)                          (the lane's two quadrature columns qx)
//   4. Y1[qx][j]  = sum_qz X2[qz][qx] S[qz][j] A = X2 (register qz),  B = S columns
//   5. Y[j][i]    = sum_qx S[qx][i] Y1[qx][j]  A = S columns,        B = Y1 (register qx)
// and Y lands as u did: register i of lane (b, j) is dof (j, i). The
// quadrature index runs over two tiles of four (NQ = 2k+1 <= 8 points,
// padded with zero table entries).
#pragma once
#include "common.h"

namespace atmo {

typedef float mx_f4 __attribute__((ext_vector_type(4)));

// 16 cells' 4x4 outer products: D += A (lane 4b+m) * B (lane 4b+n).
__device__ inline mx_f4 mx_outer(float a, float b, mx_f4 d) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, d, 0, 0, 0);
}

__device__ inline mx_f4 mx_zero() {
  mx_f4 z = {0.f, 0.f, 0.f, 0.f};
  return z;
}

// Per-lane table constants (l4 = lane % 4): Srow[t][i] = S[4t + l4][i] and
// Scol[q] = S[q][l4], zero outside the NQ x ND table.
template <int K>
struct MxTab {
  float Srow[2][4];
  float Scol[8];
  __device__ explicit MxTab(int l4) {
    constexpr int ND = K + 1, NQ = 2 * K + 1;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int q = 4 * t + l4;
        Srow[t][i] = (q < NQ && i < ND) ? (float)cT<K>.S[q < NQ ? q : 0][i < ND ? i : 0] : 0.f;
      }
#pragma unroll
    for (int q = 0; q < 8; ++q) Scol[q] = (q < NQ && l4 < ND) ? (float)cT<K>.S[q < NQ ? q : 0][l4 < ND ? l4 : 0] : 0.f;
  }
};

// Steps 1-2 for NC fields at once: values at the quadrature points,
// X2[c][tz][tx][r] = u_c(qz = 4tz + r, qx = 4tx + l4) of the lane's cell. The
// fields' (and tiles') accumulation chains are interleaved, so consecutive
// MFMAs do not wait for each other's results; each chain sums in the order
// of the vector-ALU sweep (over i, then over j).
template <int K, int NC>
__device__ inline void mx_interp(const float (&u)[NC][4], const MxTab<K>& T, mx_f4 (&X2)[NC][2][2]) {
  constexpr int ND = K + 1;
  mx_f4 X1[NC][2];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int tx = 0; tx < 2; ++tx) X1[c][tx] = mx_zero();
#pragma unroll
  for (int i = 0; i < ND; ++i)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int tx = 0; tx < 2; ++tx) X1[c][tx] = mx_outer(u[c][i], T.Srow[tx][i], X1[c][tx]);
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int tz = 0; tz < 2; ++tz)
#pragma unroll
      for (int tx = 0; tx < 2; ++tx) X2[c][tz][tx] = mx_zero();
#pragma unroll
  for (int j = 0; j < ND; ++j)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int tz = 0; tz < 2; ++tz)
#pragma unroll
        for (int tx = 0; tx < 2; ++tx) X2[c][tz][tx] = mx_outer(T.Srow[tz][j], X1[c][tx][j], X2[c][tz][tx]);
}

// Steps 4-5 for NC fields: the test-function integrals of the weighted
// point values W (the layout of mx_interp's X2): register i of lane (b, j)
// of Y[c] gets dof (j, i). Chains interleaved across fields and tiles; each
// sums over qz, then over qx, in increasing order.
template <int K, int NC>
__device__ inline void mx_integrate(const mx_f4 (&W)[NC][2][2], const MxTab<K>& T, mx_f4 (&Y)[NC]) {
  constexpr int NQ = 2 * K + 1;
  mx_f4 Y1[NC][2];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int tx = 0; tx < 2; ++tx) Y1[c][tx] = mx_zero();
#pragma unroll
  for (int qz = 0; qz < NQ; ++qz)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int tx = 0; tx < 2; ++tx) Y1[c][tx] = mx_outer(W[c][qz >> 2][tx][qz & 3], T.Scol[qz], Y1[c][tx]);
#pragma unroll
  for (int c = 0; c < NC; ++c) Y[c] = mx_zero();
#pragma unroll
  for (int qx = 0; qx < NQ; ++qx)
#pragma unroll
    for (int c = 0; c < NC; ++c) Y[c] = mx_outer(T.Scol[qx], Y1[c][qx >> 2][qx & 3], Y[c]);
}

// The mass-type apply of NC fields with one set of weights: Y[c] = M_w u_c
// (steps 1-5).
template <int K, int NC>
__device__ inline void mx_mass(const float (&u)[NC][4], const mx_f4 (&w)[2][2], const MxTab<K>& T, mx_f4 (&Y)[NC]) {
  mx_f4 X2[NC][2][2];
  mx_interp<K, NC>(u, T, X2);
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int tz = 0; tz < 2; ++tz)
#pragma unroll
      for (int tx = 0; tx < 2; ++tx) X2[c][tz][tx] *= w[tz][tx];
  mx_integrate<K, NC>(X2, T, Y);
}

}  // namespace atmo
