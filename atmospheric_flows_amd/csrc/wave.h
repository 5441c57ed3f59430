// Wave-level building blocks shared by the operator, Krylov and multigrid
// kernels: the sum-factorized mass-type apply of CPW cells held in LDS.
#pragma once
#include "common.h"

namespace atmo {

// Barrier of one wavefront: its LDS writes are visible to all its lanes after it.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-lane table columns for a lane-dependent dof index m: Si[c] = S[c][m],
// Di[c] = D[c][m] (constant-memory loads of the transposed tables).
template <int K, typename T>
__device__ inline void lane_cols(int m, T (&Si)[Cfg<K>::NQ], T (&Di)[Cfg<K>::NQ]) {
#pragma unroll
  for (int c = 0; c < Cfg<K>::NQ; ++c) {
    Si[c] = (T)cT<K>.St[m][c];
    Di[c] = (T)cT<K>.Dt[m][c];
  }
}
template <int K, typename T>
__device__ inline void lane_cols_S(int m, T (&Si)[Cfg<K>::NQ]) {
#pragma unroll
  for (int c = 0; c < Cfg<K>::NQ; ++c) Si[c] = (T)cT<K>.St[m][c];
}

// Quadrature weights of the lane's column: wq[r] = JxW(c, r) * coeff, times
// the interpolated density rho(c, r) when W (the A operator); v the cell's vertices.
template <int K, typename T, bool W>
__device__ inline void column_weights(const CellVerts& v, int c, int cw, const T* s_r, T coeff,
                                      T (&wq)[Cfg<K>::NQ], const T (&Sc)[Cfg<K>::ND]) {
  constexpr int ND = Cfg<K>::ND, NQ = Cfg<K>::NQ, ND2 = Cfg<K>::ND2;
  constexpr auto& tb = basis::kT<K>;
  const ColGeo<T> g = col_geo<K, T>(v, c);
  const T jc = g.dx * g.ze * g.wc * coeff;
#pragma unroll
  for (int r = 0; r < NQ; ++r) wq[r] = jc * (T)tb.gw[r];
  if (W) {
    T tv[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      T s = 0;
#pragma unroll
      for (int i = 0; i < ND; ++i) s += Sc[i] * s_r[cw * ND2 + j * ND + i];
      tv[j] = s;
    }
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < ND; ++j) s += (T)tb.S[r][j] * tv[j];
      wq[r] *= s;
    }
  }
}

// What a lane of the mass-type apply loads from memory besides the staged
// fields: its table row Sc and its cell's vertex words. Loaded before the
// staging (wave_mass_lane), they share the staging's memory latency.
template <int K, typename T>
struct MassLane {
  T Sc[Cfg<K>::ND];
  VertsRaw vr;
};

template <int K, typename T>
__device__ inline MassLane<K, T> wave_mass_lane(const MeshDev& m, int cell0) {
  constexpr int NQ = Cfg<K>::NQ, CPW = Cfg<K>::CPW;
  const int lane = threadIdx.x & 63;
  const int cw = lane / NQ, c = lane - cw * NQ;
  const int cell = cell0 + cw;
  MassLane<K, T> p;
  T Dc[Cfg<K>::ND];
  lane_rows<K, T>(c, p.Sc, Dc);
  p.vr = cell_verts_load(m, cw < CPW && cell < m.ncell ? cell : 0);
  return p;
}

// Column phase of the mass-type apply for the CPW cells starting at cell0:
// s_u (LDS, [NC][CPW*ND2]) -> s_red (LDS, [NC][CPW][ND][NQ]) holding
// Tv[j](c) = sum_r S[r][j] w(c,r) u(c,r). Callers barrier before reading s_red.
// pre: the lane's loads issued earlier by wave_mass_lane (else loaded here).
template <int K, typename T, int NC, bool W>
__device__ inline void wave_mass_core(const MeshDev& m, int cell0, const T* s_u, const T* s_r, T coeff, T* s_red,
                                      const MassLane<K, T>* pre = nullptr) {
  using C = Cfg<K>;
  constexpr int ND = C::ND, NQ = C::NQ, ND2 = C::ND2, CPW = C::CPW;
  constexpr auto& tb = basis::kT<K>;
  const int lane = threadIdx.x & 63;
  const int cw = lane / NQ, c = lane - cw * NQ;
  const int cell = cell0 + cw;
  if (cw >= CPW || cell >= m.ncell) return;
  T Sc[ND];
  CellVerts v;
  if (pre) {
#pragma unroll
    for (int i = 0; i < ND; ++i) Sc[i] = pre->Sc[i];
    v = cell_verts_form(m, cell, pre->vr);
  } else {
    T Dc[ND];
    lane_rows<K, T>(c, Sc, Dc);
    v = cell_verts(m, cell);
  }
  T wq[NQ];
  column_weights<K, T, W>(v, c, cw, s_r, coeff, wq, Sc);
#pragma unroll
  for (int comp = 0; comp < NC; ++comp) {
    const T* u = s_u + comp * CPW * ND2 + cw * ND2;
    T tv[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      T s = 0;
#pragma unroll
      for (int i = 0; i < ND; ++i) s += Sc[i] * u[j * ND + i];
      tv[j] = s;
    }
    T Tv[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) Tv[j] = 0;
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      T val = 0;
#pragma unroll
      for (int j = 0; j < ND; ++j) val += (T)tb.S[r][j] * tv[j];
      val *= wq[r];
#pragma unroll
      for (int j = 0; j < ND; ++j) Tv[j] += (T)tb.S[r][j] * val;
    }
    T* red = s_red + comp * CPW * ND * NQ + cw * ND * NQ;
#pragma unroll
    for (int j = 0; j < ND; ++j) red[j * NQ + c] = Tv[j];
  }
}

// Output phase: value of dof t (= cw*ND2 + jj*ND + ii) of one component;
// Si = S[.][ii] (lane_cols_S of the dof's x node).
template <int K, typename T>
__device__ inline T wave_mass_out(const T* s_red_comp, int t, const T (&Si)[Cfg<K>::NQ]) {
  constexpr int ND = Cfg<K>::ND, NQ = Cfg<K>::NQ, ND2 = Cfg<K>::ND2;
  const int cw = t / ND2, o = t - cw * ND2, jj = o / ND;
  const T* red = s_red_comp + cw * ND * NQ + jj * NQ;
  T acc = 0;
#pragma unroll
  for (int c = 0; c < NQ; ++c) acc += Si[c] * red[c];
  return acc;
}

// Eight per-lane values summed over the wave by a transposing butterfly
// (10 shuffles for the eight sums, fixed order): every lane returns the sum
// of slot (lane >> 3) & 7.
__device__ inline double wave_sum8(const double (&v)[8]) {
  const int l = threadIdx.x & 63;
  const bool b5 = l & 32, b4 = l & 16, b3 = l & 8;
  double u[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double keep = b5 ? v[4 + i] : v[i], give = b5 ? v[i] : v[4 + i];
    u[i] = keep + __shfl_xor(give, 32, 64);
  }
  double w[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double keep = b4 ? u[2 + i] : u[i], give = b4 ? u[i] : u[2 + i];
    w[i] = keep + __shfl_xor(give, 16, 64);
  }
  double x = (b3 ? w[1] : w[0]) + __shfl_xor(b3 ? w[0] : w[1], 8, 64);
  x += __shfl_xor(x, 4, 64);
  x += __shfl_xor(x, 2, 64);
  x += __shfl_xor(x, 1, 64);
  return x;
}

}  // namespace atmo
