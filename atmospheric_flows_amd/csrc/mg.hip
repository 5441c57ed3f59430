// Multigrid kernels. Between structured levels child (a,b) of coarse cell
// (I,J) is fine cell (2I+a, 2J+b); between levels of a locally refined mesh
// the coarse level's child map and the fine level's parent map say which
// cells were merged (DG embedding) and which are the same cell (identity).
#include "mg.h"
#include "mfma.h"
#include "wave.h"

namespace atmo {

// The 1D transfer matrices in constant memory: a lane's rows are read with
// lane-dependent indices. (Selected from literals by a runtime index, the 32
// FP64 entries per matrix become SGPR pairs and spill.)
template <int K>
__constant__ basis::Transfer<K> cTr = basis::make_transfer<K>();

// one thread per coarse dof
template <int K, typename T>
__global__ void k_restrict(MeshDev mf, MeshDev mc, const T* __restrict__ fine, T* __restrict__ coarse, int ncomp,
                           long Nf, long Nc) {
  constexpr int ND = K + 1, ND2 = ND * ND;
  const long tot = (long)ncomp * Nc;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int comp = (int)(t / Nc);
    const long r = t - comp * Nc;
    const int cc = (int)(r / ND2), o = (int)(r - (long)cc * ND2), m = o / ND, n = o - m * ND;
    if (mc.child && mc.child[4 * cc + 1] < 0) {
      coarse[t] = fine[comp * Nf + (long)mc.child[4 * cc] * ND2 + o];
      continue;
    }
    const int I = mc.child ? 0 : cc % mc.nx, J = mc.child ? 0 : cc / mc.nx;
    double acc = 0;
    // FP64: one row of children at a time keeps the loads in flight within the register budget
#pragma unroll(sizeof(T) == 8 ? 1 : 2)
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const long cf = mc.child ? (long)mc.child[4 * cc + a + 2 * b] : (long)(2 * J + b) * mf.nx + 2 * I + a;
        const T* u = fine + comp * Nf + cf * ND2;
#pragma unroll
        for (int jj = 0; jj < ND; ++jj) {
          const double pb = cTr<K>.P[b][jj][m];
#pragma unroll
          for (int ii = 0; ii < ND; ++ii) acc += pb * cTr<K>.P[a][ii][n] * (double)u[jj * ND + ii];
        }
      }
    coarse[t] = (T)acc;
  }
}

template <int K>
__global__ void k_project_down(MeshDev mf, MeshDev mc, const double* __restrict__ fine, double* __restrict__ coarse) {
  constexpr int ND = K + 1, ND2 = ND * ND;
  const long tot = (long)mc.ncell * ND2;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
    const int cc = (int)(t / ND2), o = (int)(t - (long)cc * ND2), m = o / ND, n = o - m * ND;
    if (mc.child && mc.child[4 * cc + 1] < 0) {
      coarse[t] = fine[(long)mc.child[4 * cc] * ND2 + o];
      continue;
    }
    const int I = mc.child ? 0 : cc % mc.nx, J = mc.child ? 0 : cc / mc.nx;
    double acc = 0;
#pragma unroll 1
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const long cf = mc.child ? (long)mc.child[4 * cc + a + 2 * b] : (long)(2 * J + b) * mf.nx + 2 * I + a;
        const double* u = fine + cf * ND2;
#pragma unroll
        for (int jj = 0; jj < ND; ++jj) {
          const double rb = cTr<K>.R[b][m][jj];
#pragma unroll
          for (int ii = 0; ii < ND; ++ii) acc += rb * cTr<K>.R[a][n][ii] * u[jj * ND + ii];
        }
      }
    coarse[t] = acc;
  }
}

static int grid_for(long n) {
  long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

template <int K, typename T>
void launch_restrict(const MeshDev& mf, const MeshDev& mc, const T* fine, T* coarse, int ncomp, long Nf, long Nc,
                     hipStream_t s) {
  hipLaunchKernelGGL((k_restrict<K, T>), dim3(grid_for(ncomp * Nc)), dim3(256), 0, s, mf, mc, fine, coarse, ncomp,
                     Nf, Nc);
  ATMO_HIP_CHECK(hipGetLastError());
}

template <int K>
void launch_project_down(const MeshDev& mf, const MeshDev& mc, const double* fine, double* coarse, hipStream_t s) {
  hipLaunchKernelGGL((k_project_down<K>), dim3(grid_for((long)mc.ncell * (K + 1) * (K + 1))), dim3(256), 0, s, mf,
                     mc, fine, coarse);
  ATMO_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------ fused smoothing
// One Chebyshev sweep of a level (all its steps) for the CPW cells of a
// one-wave block; the mass-type operators are cell-local, so x stays in LDS
// and d, b, P in registers between the steps. Step s: d = f1[s] d +
// f2[s] P (b - A x), x += d; with ZG (zero initial guess) step 0 is
// d = x = P b f2[0] without an operator application. PRO: the sweep starts
// from x + prolongation of the coarse level's xc (the V-cycle's coarse-grid
// correction). RES: the sweep ends with res = b - A x (the V-cycle's
// residual for the restriction). Same expressions, in the same order, as
// the first-step update, the fused operator-and-update step, the DG
// embedding prolongation and the residual of the two-kernel forms, so the
// values are those of those kernels.
// The FP32 sweeps are compiled for 5 waves per SIMD (they fit 102 VGPRs
// without spills; the compiler's default choice, 118, allowed 4).
template <int K, typename T, int NC, bool W, bool ZG, bool PRO, bool RES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 5 : 1))) void k_cheb_sweep(MeshDev m, MeshDev mc, T* __restrict__ x, const T* __restrict__ xc,
                                                   const T* __restrict__ rho, T coeff, const T* __restrict__ b,
                                                   const T* __restrict__ P, ChebSteps cs, T* __restrict__ res,
                                                   long Ns, long Nc) {
  using C = Cfg<K>;
  constexpr int ND = C::ND, NQ = C::NQ, ND2 = C::ND2, CPW = C::CPW, NPASS = (CPW * ND2 + 63) / 64;
  constexpr auto& tb = basis::kT<K>;
  constexpr int NS = NC + (W ? 1 : 0);
  __shared__ T s_u[NS][CPW * ND2];  // x (components), then rho when W
  __shared__ T s_red[NC][CPW][ND][NQ];
  const int blk = cell_block();
  const long base = (long)blk * CPW * ND2;
  const long lim = (long)m.ncell * ND2;
  // the lane's table row and cell vertex words, loaded with the staged fields
  const MassLane<K, T> lane_pre = wave_mass_lane<K, T>(m, blk * CPW);
  // stage x (unless the guess is zero) and rho
  {
    const T* srcs[NS];
    int nsrc = 0;
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) srcs[comp] = x + comp * Ns;
    if constexpr (W) srcs[NS - 1] = rho;
    if constexpr (!ZG) {
      stage_fields<K, NS, T, T>(&s_u[0][0], srcs, m.ncell);
    } else if constexpr (W) {
      const T* s1[1] = {rho};
      stage_fields<K, 1, T, T>(&s_u[NS - 1][0], s1, m.ncell);
    }
    (void)nsrc;
  }
  // per-dof b and P in registers
  T bv[NPASS][NC], pv[NPASS][NC], dv[NPASS][NC];
  bool ok[NPASS];
#pragma unroll
  for (int ps = 0; ps < NPASS; ++ps) {
    const int t = threadIdx.x + 64 * ps;
    ok[ps] = t < CPW * ND2 && base + t < lim;
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) {
      const long k = comp * Ns + (ok[ps] ? base + t : 0);
      bv[ps][comp] = ok[ps] ? b[k] : (T)0;
      pv[ps][comp] = ok[ps] ? P[k] : (T)0;
      dv[ps][comp] = (T)0;
    }
  }
  __syncthreads();
  if constexpr (PRO) {
    // x += prolongation of the parent cell's coarse values (DG embedding), or the identity
#pragma unroll
    for (int ps = 0; ps < NPASS; ++ps) {
      const int t = threadIdx.x + 64 * ps;
      if (!ok[ps]) continue;
      const long r = base + t;
      const int cf = (int)(r / ND2), o = (int)(r - (long)cf * ND2), jj = o / ND, ii = o - jj * ND;
      int cc, a = 0, bb = 0;
      bool ident = false;
      if (m.parent) {
        cc = m.parent[2 * cf];
        const int code = m.parent[2 * cf + 1];
        if (code < 0) ident = true;
        else {
          a = code & 1;
          bb = code >> 1;
        }
      } else {
        const int i = cf % m.nx, j = cf / m.nx;
        cc = (j >> 1) * mc.nx + (i >> 1);
        a = i & 1;
        bb = j & 1;
      }
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) {
        const T* u = xc + comp * Nc + (long)cc * ND2;
        if (ident) {
          s_u[comp][t] += u[o];
        } else {
          const double* Pa = cTr<K>.P[a][ii];
          const double* Pb = cTr<K>.P[bb][jj];
          double acc = 0;
#pragma unroll
          for (int mm = 0; mm < ND; ++mm) {
            const double pb = Pb[mm];
#pragma unroll
            for (int n = 0; n < ND; ++n) acc += pb * Pa[n] * (double)u[mm * ND + n];
          }
          s_u[comp][t] += (T)acc;
        }
      }
    }
    __syncthreads();
  }
  // column weights once per lane
  const int lane = threadIdx.x & 63;
  const int cw = lane / NQ, c = lane - cw * NQ;
  const int cell = blk * CPW + cw;
  const bool col_active = cw < CPW && cell < m.ncell;
  T Sc[ND], wq[NQ];
#pragma unroll
  for (int i = 0; i < ND; ++i) Sc[i] = lane_pre.Sc[i];
  if (col_active) column_weights<K, T, W>(cell_verts_form(m, cell, lane_pre.vr), c, cw, &s_u[NS - 1][0], coeff, wq, Sc);
  T Si[NPASS][NQ];
#pragma unroll
  for (int ps = 0; ps < NPASS; ++ps) lane_cols_S<K, T>((threadIdx.x + 64 * ps) % ND, Si[ps]);
  // A x of the staged x into s_red (the column phase of wave_mass_core with the cached weights)
  auto apply = [&]() {
    if (col_active) {
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) {
        const T* u = &s_u[comp][cw * ND2];
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
#pragma unroll
        for (int j = 0; j < ND; ++j) s_red[comp][cw][j][c] = Tv[j];
      }
    }
    __syncthreads();
  };
  for (int s = 0; s < cs.n; ++s) {
    const T f1 = (T)cs.f1[s], f2 = (T)cs.f2[s];
    if (ZG && s == 0) {
#pragma unroll
      for (int ps = 0; ps < NPASS; ++ps) {
        const int t = threadIdx.x + 64 * ps;
        if (!ok[ps]) continue;
#pragma unroll
        for (int comp = 0; comp < NC; ++comp) {
          const T v = pv[ps][comp] * bv[ps][comp] * f2;
          dv[ps][comp] = v;
          s_u[comp][t] = v;
        }
      }
      __syncthreads();
      continue;
    }
    apply();
#pragma unroll
    for (int ps = 0; ps < NPASS; ++ps) {
      const int t = threadIdx.x + 64 * ps;
      if (!ok[ps]) continue;
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) {
        const T ax = wave_mass_out<K, T>(&s_red[comp][0][0][0], t, Si[ps]);
        const T v = f1 * dv[ps][comp] + f2 * pv[ps][comp] * (bv[ps][comp] - ax);
        dv[ps][comp] = v;
        s_u[comp][t] = s_u[comp][t] + v;
      }
    }
    __syncthreads();
  }
  if constexpr (RES) {
    apply();
#pragma unroll
    for (int ps = 0; ps < NPASS; ++ps) {
      const int t = threadIdx.x + 64 * ps;
      if (!ok[ps]) continue;
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) {
        const T ax = wave_mass_out<K, T>(&s_red[comp][0][0][0], t, Si[ps]);
        res[comp * Ns + base + t] = -ax + bv[ps][comp];
      }
    }
  }
#pragma unroll
  for (int ps = 0; ps < NPASS; ++ps) {
    const int t = threadIdx.x + 64 * ps;
    if (!ok[ps]) continue;
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) x[comp * Ns + base + t] = s_u[comp][t];
  }
}

// The same sweep on the matrix cores (FP32, k = 2 and 3): 16 cells per wave,
// four lanes per cell, every vector of the sweep (x, d, b, P) and the
// quadrature weights in registers in the layout of mfma.h, each operator
// application 33 (k = 2) or 45 (k = 3) MFMA outer-product steps with no LDS
// and no barrier. Same steps and update expressions as k_cheb_sweep; every
// MFMA step is one fused multiply-add per entry and each contraction is
// accumulated in k_cheb_sweep's order, so the values are bitwise those of
// k_cheb_sweep.
constexpr int MX_CELLS = 16;

template <int K, int NC, bool W, bool ZG, bool PRO, bool RES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_cheb_sweep_mx(MeshDev m, MeshDev mc, float* __restrict__ x,
                                                      const float* __restrict__ xc, const float* __restrict__ rho,
                                                      float coeff, const float* __restrict__ b,
                                                      const float* __restrict__ P, ChebSteps cs,
                                                      float* __restrict__ res, long Ns, long Nc) {
  static_assert(K == 2 || K == 3, "matrix-core sweep: k = 2 or 3");
  constexpr int ND = K + 1, NQ = 2 * K + 1, ND2 = ND * ND;
  const int lane = threadIdx.x & 63, l4 = lane & 3;
  const int cell = cell_block() * MX_CELLS + (lane >> 2);
  const bool live = cell < m.ncell && l4 < ND;  // the lane's dof row exists
  const long row = live ? (long)cell * ND2 + l4 * ND : 0;
  const VertsRaw vr = cell_verts_load(m, cell < m.ncell ? cell : 0);
  float xv[NC][4], bv[NC][4], pv[NC][4], dv[NC][4], rv[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const bool on = live && i < ND;
    const long k = row + (i < ND ? i : 0);
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) {
      xv[comp][i] = (!ZG && on) ? x[comp * Ns + k] : 0.f;
      bv[comp][i] = on ? b[comp * Ns + k] : 0.f;
      pv[comp][i] = on ? P[comp * Ns + k] : 0.f;
      dv[comp][i] = 0.f;
    }
    rv[i] = (W && on) ? rho[k] : 0.f;
  }
  const MxTab<K> T(l4);
  if constexpr (PRO) {
    // x += prolongation of the parent cell's coarse values (DG embedding), or the identity; the expressions
    // of k_cheb_sweep
    if (live) {
      int cc, a = 0, bb = 0;
      bool ident = false;
      if (m.parent) {
        cc = m.parent[2 * cell];
        const int code = m.parent[2 * cell + 1];
        if (code < 0) ident = true;
        else {
          a = code & 1;
          bb = code >> 1;
        }
      } else {
        const int i = cell % m.nx, j = cell / m.nx;
        cc = (j >> 1) * mc.nx + (i >> 1);
        a = i & 1;
        bb = j & 1;
      }
      const int jj = l4;
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) {
        const float* u = xc + comp * Nc + (long)cc * ND2;
#pragma unroll
        for (int ii = 0; ii < ND; ++ii) {
          if (ident) {
            xv[comp][ii] += u[jj * ND + ii];
          } else {
            const double* Pa = cTr<K>.P[a][ii];
            const double* Pb = cTr<K>.P[bb][jj];
            double acc = 0;
#pragma unroll
            for (int mm = 0; mm < ND; ++mm) {
              const double pb = Pb[mm];
#pragma unroll
              for (int n = 0; n < ND; ++n) acc += pb * Pa[n] * (double)u[mm * ND + n];
            }
            xv[comp][ii] += (float)acc;
          }
        }
      }
    }
  }
  // quadrature weights of the lane's points (qz = 4tz + r, qx = 4tx + l4): JxW * coeff (column_weights'
  // expressions), times the density there when W; zero at the padding points
  mx_f4 w[2][2];
  {
    float rv1[1][4] = {{rv[0], rv[1], rv[2], rv[3]}};
    mx_f4 R[1][2][2];
    if constexpr (W) mx_interp<K, 1>(rv1, T, R);
    const CellVerts v = cell_verts_form(m, cell < m.ncell ? cell : 0, vr);
#pragma unroll
    for (int tx = 0; tx < 2; ++tx) {
      const int qx = 4 * tx + l4;
      float jc = 0.f;
      if (qx < NQ) {
        const ColGeo<float> g = col_geo<K, float>(v, qx);
        jc = g.dx * g.ze * g.wc * coeff;
      }
#pragma unroll
      for (int tz = 0; tz < 2; ++tz)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int qz = 4 * tz + r;
          float wq = qz < NQ ? jc * (float)cT<K>.gw[qz < NQ ? qz : 0] : 0.f;
          if constexpr (W) wq *= R[0][tz][tx][r];
          w[tz][tx][r] = wq;
        }
    }
  }
  for (int s = 0; s < cs.n; ++s) {
    const float f1 = (float)cs.f1[s], f2 = (float)cs.f2[s];
    if (ZG && s == 0) {
#pragma unroll
      for (int comp = 0; comp < NC; ++comp)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float v = pv[comp][i] * bv[comp][i] * f2;
          dv[comp][i] = v;
          xv[comp][i] = v;
        }
      continue;
    }
    mx_f4 ax[NC];
    mx_mass<K, NC>(xv, w, T, ax);
#pragma unroll
    for (int comp = 0; comp < NC; ++comp)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float v = f1 * dv[comp][i] + f2 * pv[comp][i] * (bv[comp][i] - ax[comp][i]);
        dv[comp][i] = v;
        xv[comp][i] = xv[comp][i] + v;
      }
  }
  if constexpr (RES) {
    mx_f4 ax[NC];
    mx_mass<K, NC>(xv, w, T, ax);
    if (live)
#pragma unroll
      for (int comp = 0; comp < NC; ++comp)
#pragma unroll
        for (int i = 0; i < ND; ++i) res[comp * Ns + row + i] = -ax[comp][i] + bv[comp][i];
  }
  if (live)
#pragma unroll
    for (int comp = 0; comp < NC; ++comp)
#pragma unroll
      for (int i = 0; i < ND; ++i) x[comp * Ns + row + i] = xv[comp][i];
}

template <int K, bool ZG, bool PRO, bool RES>
static void sweep_launch_mx(const MeshDev& m, const MeshDev& mc, int ncomp, bool weighted, float* x,
                            const float* xc, const float* rho, float coeff, const float* b, const float* P,
                            const ChebSteps& cs, float* res, long Ns, long Nc, hipStream_t s) {
  const int grid = (m.ncell + MX_CELLS - 1) / MX_CELLS;
  if (weighted) {
    if (ncomp != 2) throw std::invalid_argument("weighted mass operator expects 2 components");
    hipLaunchKernelGGL((k_cheb_sweep_mx<K, 2, true, ZG, PRO, RES>), dim3(grid), dim3(64), 0, s, m, mc, x, xc, rho,
                       coeff, b, P, cs, res, Ns, Nc);
  } else if (ncomp == 1) {
    hipLaunchKernelGGL((k_cheb_sweep_mx<K, 1, false, ZG, PRO, RES>), dim3(grid), dim3(64), 0, s, m, mc, x, xc, rho,
                       coeff, b, P, cs, res, Ns, Nc);
  } else {
    hipLaunchKernelGGL((k_cheb_sweep_mx<K, 2, false, ZG, PRO, RES>), dim3(grid), dim3(64), 0, s, m, mc, x, xc, rho,
                       coeff, b, P, cs, res, Ns, Nc);
  }
  ATMO_HIP_CHECK(hipGetLastError());
}

template <int K, typename T, bool ZG, bool PRO, bool RES>
static void sweep_launch(const MeshDev& m, const MeshDev& mc, int ncomp, bool weighted, T* x, const T* xc,
                         const T* rho, T coeff, const T* b, const T* P, const ChebSteps& cs, T* res, long Ns, long Nc,
                         bool matrix_cores, hipStream_t s) {
  const int grid = (m.ncell + Cfg<K>::CPW - 1) / Cfg<K>::CPW;
  if (grid == 0) return;
  if constexpr (std::is_same<T, float>::value && (K == 2 || K == 3)) {
    if (matrix_cores) {
      sweep_launch_mx<K, ZG, PRO, RES>(m, mc, ncomp, weighted, x, xc, rho, coeff, b, P, cs, res, Ns, Nc, s);
      return;
    }
  }
  (void)matrix_cores;
  if (weighted) {
    if (ncomp != 2) throw std::invalid_argument("weighted mass operator expects 2 components");
    hipLaunchKernelGGL((k_cheb_sweep<K, T, 2, true, ZG, PRO, RES>), dim3(grid), dim3(64), 0, s, m, mc, x, xc, rho,
                       coeff, b, P, cs, res, Ns, Nc);
  } else if (ncomp == 1) {
    hipLaunchKernelGGL((k_cheb_sweep<K, T, 1, false, ZG, PRO, RES>), dim3(grid), dim3(64), 0, s, m, mc, x, xc, rho,
                       coeff, b, P, cs, res, Ns, Nc);
  } else {
    hipLaunchKernelGGL((k_cheb_sweep<K, T, 2, false, ZG, PRO, RES>), dim3(grid), dim3(64), 0, s, m, mc, x, xc, rho,
                       coeff, b, P, cs, res, Ns, Nc);
  }
  ATMO_HIP_CHECK(hipGetLastError());
}

template <int K, typename T>
void launch_presmooth_residual(const MeshDev& m, int ncomp, bool weighted, T* x, const T* rho, T coeff, const T* b,
                               const T* P, const ChebSteps& cs, T* res, long Ns, bool matrix_cores, hipStream_t s) {
  if (cs.n < 1 || cs.n > ChebSteps::MAX) throw std::invalid_argument("Chebyshev degree out of range");
  sweep_launch<K, T, true, false, true>(m, m, ncomp, weighted, x, nullptr, rho, coeff, b, P, cs, res, Ns, 0,
                                        matrix_cores, s);
}

template <int K, typename T>
void launch_prolongate_postsmooth(const MeshDev& m, const MeshDev& mc, int ncomp, bool weighted, T* x, const T* xc,
                                  const T* rho, T coeff, const T* b, const T* P, const ChebSteps& cs, long Ns,
                                  long Nc, bool matrix_cores, hipStream_t s) {
  if (cs.n < 1 || cs.n > ChebSteps::MAX) throw std::invalid_argument("Chebyshev degree out of range");
  sweep_launch<K, T, false, true, false>(m, mc, ncomp, weighted, x, xc, rho, coeff, b, P, cs, nullptr, Ns, Nc,
                                         matrix_cores, s);
}

// ------------------------------------------------------------ coarse CG
__device__ inline double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;  // all lanes hold the sum
}

template <int K, int NC, bool W>
__device__ void coarse_apply(const MeshDev& m, const double* rho, double coeff, const double* src, double* dst,
                             long Ns, double* s_u, double* s_r, double* s_red) {
  using C = Cfg<K>;
  constexpr int ND2 = C::ND2, CPW = C::CPW;
  const int passes = (m.ncell + CPW - 1) / CPW;
  for (int p = 0; p < passes; ++p) {
    const int cell0 = p * CPW;
    const long base = (long)cell0 * ND2, lim = (long)m.ncell * ND2;
    for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
      const long gi = base + t;
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) s_u[comp * CPW * ND2 + t] = gi < lim ? src[comp * Ns + gi] : 0.0;
      if (W) s_r[t] = gi < lim ? rho[gi] : 0.0;
    }
    __syncthreads();
    wave_mass_core<K, double, NC, W>(m, cell0, s_u, s_r, coeff, s_red);
    __syncthreads();
    double Si[C::NQ];
    lane_cols_S<K, double>(threadIdx.x % C::ND, Si);
    for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
      const long gi = base + t;
      if (gi >= lim) break;
      if (64 % C::ND != 0) lane_cols_S<K, double>(t % C::ND, Si);
#pragma unroll
      for (int comp = 0; comp < NC; ++comp)
        dst[comp * Ns + gi] = wave_mass_out<K, double>(s_red + comp * CPW * C::ND * C::NQ, t, Si);
    }
    __syncthreads();
  }
}

template <int K, int NC, bool W>
__global__ __launch_bounds__(64) void k_coarse_cg(MeshDev m, const double* __restrict__ rho, double coeff,
                                                  const double* __restrict__ b, double* __restrict__ x,
                                                  double* __restrict__ g, double* __restrict__ d,
                                                  double* __restrict__ h, const double* tolp, int max_its, int* its,
                                                  long Ns) {
  using C = Cfg<K>;
  __shared__ double s_u[NC * C::CPW * C::ND2];
  __shared__ double s_r[W ? C::CPW * C::ND2 : 1];
  __shared__ double s_red[NC * C::CPW * C::ND * C::NQ];
  const long n = (long)NC * Ns;
  const double tol = *tolp;
  double gg = 0;
  for (long i = threadIdx.x; i < n; i += 64) {
    const double gi = -b[i];
    x[i] = 0.0;
    g[i] = gi;
    d[i] = -gi;
    gg += gi * gi;
  }
  gg = wsum(gg);
  double res = sqrt(gg);
  int it = 0;
  int status = 0;
  if (res <= tol) status = 1;
  else if (max_its <= 0) status = -1;
  double gh = gg;
  __syncthreads();
  while (status == 0) {
    ++it;
    coarse_apply<K, NC, W>(m, rho, coeff, d, h, Ns, s_u, s_r, s_red);
    double dh = 0;
    for (long i = threadIdx.x; i < n; i += 64) dh += d[i] * h[i];
    dh = wsum(dh);
    const double alpha = gh / dh;
    gg = 0;
    for (long i = threadIdx.x; i < n; i += 64) {
      x[i] += alpha * d[i];
      const double gi = g[i] + alpha * h[i];
      g[i] = gi;
      gg += gi * gi;
    }
    gg = wsum(gg);
    res = sqrt(gg);
    if (res <= tol) { status = 1; break; }
    if (it >= max_its || isnan(res)) { status = -1; break; }
    const double beta = gg / gh;
    gh = gg;
    for (long i = threadIdx.x; i < n; i += 64) d[i] = -g[i] + beta * d[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) *its = status == 1 ? it : -1;
}

// The same CG with every vector of the level (x, g, d, h and the density)
// held in LDS for the whole solve, and each lane's column weights (the
// Jacobian, quadrature weights and interpolated density of its quadrature
// column in each pass, which do not change across iterations) formed once
// into LDS: an iteration makes no global memory access. Same operations in
// the same order as k_coarse_cg (one wave; the operator is applied one
// component at a time with wave_mass_core's arithmetic), so the results are
// those of k_coarse_cg.
constexpr size_t COARSE_LDS_MAX = 60 * 1024;  // bytes of level vectors and weights a coarse solve keeps in LDS

template <int K>
size_t coarse_lds_bytes(int ncell, int ncomp, bool weighted, long Ns) {
  const int passes = (ncell + Cfg<K>::CPW - 1) / Cfg<K>::CPW;
  return ((size_t)(4 * ncomp + (weighted ? 1 : 0)) * Ns + (size_t)passes * 64 * Cfg<K>::NQ) * sizeof(double);
}

// TB, TX: the types b is read and x written in (float in the mixed form, whose levels are FP32: the
// conversions are those of a cast kernel before and after the solve).
template <int K, int NC, bool W, typename TB, typename TX>
__global__ __launch_bounds__(64) void k_coarse_cg_lds(MeshDev m, const double* __restrict__ rho, double coeff,
                                                      const TB* __restrict__ b, TX* __restrict__ x_out,
                                                      const double* tolp, int max_its, int* its, long Ns) {
  using C = Cfg<K>;
  constexpr int ND = C::ND, NQ = C::NQ, ND2 = C::ND2, CPW = C::CPW;
  constexpr auto& tb = basis::kT<K>;
  extern __shared__ double lv[];
  __shared__ double s_red[CPW * C::ND * C::NQ];
  const long n = (long)NC * Ns;
  double *x = lv, *g = lv + n, *d = lv + 2 * n, *h = lv + 3 * n, *r = lv + 4 * n;
  double* wql = lv + 4 * n + (W ? Ns : 0);  // [pass][lane][NQ]
  const double tol = *tolp;
  double gg = 0;
  for (long i = threadIdx.x; i < n; i += 64) {
    const double gi = -(double)b[i];
    x[i] = 0.0;
    g[i] = gi;
    d[i] = -gi;
    gg += gi * gi;
  }
  if (W)
    for (long i = threadIdx.x; i < Ns; i += 64) r[i] = rho[i];
  gg = wsum(gg);
  double res = sqrt(gg);
  int it = 0;
  int status = 0;
  if (res <= tol) status = 1;
  else if (max_its <= 0) status = -1;
  double gh = gg;
  wave_sync();
  const int passes = (m.ncell + CPW - 1) / CPW;
  const int lane = threadIdx.x & 63, cw = lane / NQ, c = lane - cw * NQ;
  double Sc[ND];
  {
    double Dc[ND];
    lane_rows<K, double>(c, Sc, Dc);
  }
  // the lane's column weights in every pass (wave_mass_core's column_weights, once per solve)
  for (int p = 0; p < passes; ++p) {
    const int cell = p * CPW + cw;
    if (cw < CPW && cell < m.ncell) {
      double wq[NQ];
      column_weights<K, double, W>(cell_verts(m, cell), c, cw, r + (long)p * CPW * ND2, coeff, wq, Sc);
#pragma unroll
      for (int q = 0; q < NQ; ++q) wql[((long)p * 64 + lane) * NQ + q] = wq[q];
    }
  }
  wave_sync();
  double Si[C::NQ];
  lane_cols_S<K, double>(threadIdx.x % C::ND, Si);
  while (status == 0) {
    ++it;
    // h = A d, pass by pass and component by component
    for (int p = 0; p < passes; ++p) {
      const int cell0 = p * CPW;
      const long base = (long)cell0 * ND2, lim = (long)m.ncell * ND2;
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) {
        // wave_mass_core's column phase with the stored weights
        if (cw < CPW && cell0 + cw < m.ncell) {
          const double* u = d + comp * Ns + base + cw * ND2;
          const double* wq = wql + ((long)p * 64 + lane) * NQ;
          double tv[ND];
#pragma unroll
          for (int j = 0; j < ND; ++j) {
            double sum = 0;
#pragma unroll
            for (int i = 0; i < ND; ++i) sum += Sc[i] * u[j * ND + i];
            tv[j] = sum;
          }
          double Tv[ND];
#pragma unroll
          for (int j = 0; j < ND; ++j) Tv[j] = 0;
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            double val = 0;
#pragma unroll
            for (int j = 0; j < ND; ++j) val += (double)tb.S[q][j] * tv[j];
            val *= wq[q];
#pragma unroll
            for (int j = 0; j < ND; ++j) Tv[j] += (double)tb.S[q][j] * val;
          }
#pragma unroll
          for (int j = 0; j < ND; ++j) s_red[cw * ND * NQ + j * NQ + c] = Tv[j];
        }
        wave_sync();
        for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
          if (base + t >= lim) break;
          if (64 % C::ND != 0) lane_cols_S<K, double>(t % C::ND, Si);
          h[comp * Ns + base + t] = wave_mass_out<K, double>(s_red, t, Si);
        }
        if (64 % C::ND != 0) lane_cols_S<K, double>(threadIdx.x % C::ND, Si);
        wave_sync();
      }
    }
    double dh = 0;
    for (long i = threadIdx.x; i < n; i += 64) dh += d[i] * h[i];
    dh = wsum(dh);
    const double alpha = gh / dh;
    gg = 0;
    for (long i = threadIdx.x; i < n; i += 64) {
      x[i] += alpha * d[i];
      const double gi = g[i] + alpha * h[i];
      g[i] = gi;
      gg += gi * gi;
    }
    gg = wsum(gg);
    res = sqrt(gg);
    if (res <= tol) { status = 1; break; }
    if (it >= max_its || isnan(res)) { status = -1; break; }
    const double beta = gg / gh;
    gh = gg;
    for (long i = threadIdx.x; i < n; i += 64) d[i] = -g[i] + beta * d[i];
    wave_sync();
  }
  wave_sync();
  for (long i = threadIdx.x; i < n; i += 64) x_out[i] = (TX)x[i];
  if (threadIdx.x == 0) *its = status == 1 ? it : -1;
}

template <int K, typename TB, typename TX>
static bool coarse_cg_lds(const MeshDev& m, int ncomp, bool weighted, const double* rho, double coeff, const TB* b,
                          TX* x, const double* tol, int max_its, int* its, long Ns, hipStream_t s) {
  const size_t lds = coarse_lds_bytes<K>(m.ncell, ncomp, weighted, Ns);
  if (lds > COARSE_LDS_MAX) return false;
  if (weighted)
    hipLaunchKernelGGL((k_coarse_cg_lds<K, 2, true, TB, TX>), dim3(1), dim3(64), lds, s, m, rho, coeff, b, x, tol,
                       max_its, its, Ns);
  else if (ncomp == 1)
    hipLaunchKernelGGL((k_coarse_cg_lds<K, 1, false, TB, TX>), dim3(1), dim3(64), lds, s, m, rho, coeff, b, x, tol,
                       max_its, its, Ns);
  else
    hipLaunchKernelGGL((k_coarse_cg_lds<K, 2, false, TB, TX>), dim3(1), dim3(64), lds, s, m, rho, coeff, b, x, tol,
                       max_its, its, Ns);
  ATMO_HIP_CHECK(hipGetLastError());
  return true;
}

template <int K>
void launch_coarse_cg(const MeshDev& m, int ncomp, bool weighted, const double* rho, double coeff, const double* b,
                      double* x, double* g, double* d, double* h, const double* tol, int max_its, int* its, long Ns,
                      hipStream_t s) {
  if (coarse_cg_lds<K, double, double>(m, ncomp, weighted, rho, coeff, b, x, tol, max_its, its, Ns, s)) return;
  if (weighted)
    hipLaunchKernelGGL((k_coarse_cg<K, 2, true>), dim3(1), dim3(64), 0, s, m, rho, coeff, b, x, g, d, h, tol, max_its,
                       its, Ns);
  else if (ncomp == 1)
    hipLaunchKernelGGL((k_coarse_cg<K, 1, false>), dim3(1), dim3(64), 0, s, m, rho, coeff, b, x, g, d, h, tol,
                       max_its, its, Ns);
  else
    hipLaunchKernelGGL((k_coarse_cg<K, 2, false>), dim3(1), dim3(64), 0, s, m, rho, coeff, b, x, g, d, h, tol,
                       max_its, its, Ns);
  ATMO_HIP_CHECK(hipGetLastError());
}

template <int K>
void launch_coarse_cg_f32(const MeshDev& m, int ncomp, bool weighted, const double* rho, double coeff, const float* b,
                          float* x, double* b64, double* x64, double* g, double* d, double* h, const double* tol,
                          int max_its, int* its, long Ns, hipStream_t s) {
  if (coarse_cg_lds<K, float, float>(m, ncomp, weighted, rho, coeff, b, x, tol, max_its, its, Ns, s)) return;
  const long n = (long)ncomp * Ns;
  vec_cast<double, float>(n, b, b64, s);
  launch_coarse_cg<K>(m, ncomp, weighted, rho, coeff, b64, x64, g, d, h, tol, max_its, its, Ns, s);
  vec_cast<float, double>(n, x64, x, s);
}

// CG of one level's eigenvalue estimate per workgroup (1024 threads), all iterations inside: the
// coefficients alpha_k, beta_k go to E.hist for the Lanczos tridiagonal matrix (the iterate itself is
// not formed: nothing reads it). Every sum has a fixed order: a thread takes elements tid + 1024 j in
// increasing j, then each wave's shuffle tree, then the waves in order. Iterations after the stopping
// test succeeds or fails record nothing, so the loop ends there.
template <int K, int NC, bool W>
__global__ __launch_bounds__(1024) void k_eig_cg(EigBatch B, double coeff, int nit, double tol) {
  using C = Cfg<K>;
  constexpr int ND2 = C::ND2, CPW = C::CPW, NW = 16, NT = NW * 64;
  constexpr int NU = NC * CPW * ND2, NS = NU + (W ? CPW * ND2 : 0), NLD = (NS + 63) / 64;
  constexpr int U = 4;  // element loops: batches of U elements per thread, every load issued before use
  __shared__ double s_u[NW][NC * CPW * ND2];
  __shared__ double s_r[NW][W ? CPW * ND2 : 1];
  __shared__ double s_red[NW][NC * CPW * C::ND * C::NQ];
  __shared__ double s_sum[2][NW];
  const EigLevel& E = B.L[blockIdx.x];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long n = (long)NC * E.Ns;
  // block sums of a and (two) b
  auto bsum = [&](double& a, double& b, bool two) {
    a = wsum(a);
    if (two) b = wsum(b);
    if (lane == 0) {
      s_sum[0][w] = a;
      if (two) s_sum[1][w] = b;
    }
    __syncthreads();
    double ta = 0, tb = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      ta += s_sum[0][i];
      if (two) tb += s_sum[1][i];
    }
    __syncthreads();
    a = ta;
    b = tb;
  };
  // dst = A src on the level: the waves take interleaved passes of CPW cells, each in its own LDS; the
  // next pass's values are loaded while the current pass is computed
  auto stage_load = [&](int p, const double* src, double (&v)[NLD]) {
    const long base = (long)p * CPW * ND2, lim = (long)E.m.ncell * ND2;
#pragma unroll
    for (int k = 0; k < NLD; ++k) {
      const int q = lane + 64 * k;
      const int comp = q < NU ? q / (CPW * ND2) : 0;
      const int t = q < NU ? q - comp * CPW * ND2 : q - NU;
      const long gi = base + t;
      v[k] = 0.0;
      if (q < NS && gi < lim) v[k] = q < NU ? src[comp * E.Ns + gi] : E.rho[gi];
    }
  };
  auto apply = [&](const double* src, double* dst) {
    const int passes = (E.m.ncell + CPW - 1) / CPW;
    double v[NLD];
    if (w < passes) stage_load(w, src, v);
    for (int p = w; p < passes; p += NW) {
#pragma unroll
      for (int k = 0; k < NLD; ++k) {
        const int q = lane + 64 * k;
        if (q < NU) s_u[w][q] = v[k];
        else if (W && q < NS) s_r[w][q - NU] = v[k];
      }
      if (p + NW < passes) stage_load(p + NW, src, v);
      wave_sync();
      const int cell0 = p * CPW;
      wave_mass_core<K, double, NC, W>(E.m, cell0, s_u[w], s_r[w], coeff, s_red[w]);
      wave_sync();
      const long base = (long)cell0 * ND2, lim = (long)E.m.ncell * ND2;
      double Si[C::NQ];
      lane_cols_S<K, double>(lane % C::ND, Si);
      for (int t = lane; t < CPW * ND2; t += 64) {
        const long gi = base + t;
        if (gi >= lim) break;
        if (64 % C::ND != 0) lane_cols_S<K, double>(t % C::ND, Si);
#pragma unroll
        for (int comp = 0; comp < NC; ++comp)
          dst[comp * E.Ns + gi] = wave_mass_out<K, double>(s_red[w] + comp * CPW * C::ND * C::NQ, t, Si);
      }
      wave_sync();
    }
  };
  double* __restrict__ g = E.g;
  double* __restrict__ d = E.d;
  double* __restrict__ h = E.h;
  const double* __restrict__ P = E.P;
  const double* __restrict__ b = E.b;
  double part = 0, unused = 0;
  for (long i0 = threadIdx.x; i0 < n; i0 += U * NT) {
    double bv[U], pv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = i0 + (long)u * NT;
      bv[u] = i < n ? b[i] : 0.0;
      pv[u] = i < n ? P[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = i0 + (long)u * NT;
      if (i >= n) break;
      const double gi = -bv[u];
      const double hi = gi * pv[u];
      g[i] = gi;
      h[i] = hi;
      d[i] = -hi;
      part += gi * hi;
    }
  }
  bsum(part, unused, false);
  double gh = part;
  int it = 0, done = 0;
  for (int k = 0; k < nit; ++k) {
    __syncthreads();
    apply(d, h);
    __syncthreads();
    part = 0;
    for (long i0 = threadIdx.x; i0 < n; i0 += U * NT) {
      double dv[U], hv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long i = i0 + (long)u * NT;
        dv[u] = i < n ? d[i] : 0.0;
        hv[u] = i < n ? h[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + (long)u * NT < n) part += dv[u] * hv[u];
    }
    bsum(part, unused, false);
    const double alpha = gh / part;
    if (threadIdx.x == 0) E.hist[2 * k] = alpha;
    // g += alpha h with |g|^2, and (unless this is the last iteration) h = P g with g.h
    const bool more = k + 1 < nit;
    double pgg = 0, pgh = 0;
    for (long i0 = threadIdx.x; i0 < n; i0 += U * NT) {
      double gv[U], hv[U], pv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long i = i0 + (long)u * NT;
        gv[u] = i < n ? g[i] : 0.0;
        hv[u] = i < n ? h[i] : 0.0;
        pv[u] = i < n && more ? P[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long i = i0 + (long)u * NT;
        if (i >= n) break;
        const double gi = gv[u] + alpha * hv[u];
        g[i] = gi;
        pgg += gi * gi;
        if (more) {
          const double hi = gi * pv[u];
          h[i] = hi;
          pgh += gi * hi;
        }
      }
    }
    bsum(pgg, pgh, more);
    const double res = sqrt(pgg);
    ++it;
    if (res <= tol) done = 1;
    else if (it >= nit || isnan(res)) done = -1;
    if (done) break;
    const double beta = pgh / gh;
    gh = pgh;
    if (threadIdx.x == 0) E.hist[2 * k + 1] = beta;
    for (long i0 = threadIdx.x; i0 < n; i0 += U * NT) {
      double hv[U], dv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long i = i0 + (long)u * NT;
        hv[u] = i < n ? h[i] : 0.0;
        dv[u] = i < n ? d[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long i = i0 + (long)u * NT;
        if (i < n) d[i] = -hv[u] + beta * dv[u];
      }
    }
  }
  if (threadIdx.x == 0) {
    E.st->it = it;
    E.st->done = done;
  }
}

template <int K>
void launch_eig_cg_batch(const EigBatch& B, int ncomp, bool weighted, double coeff, int nit, double tol,
                         hipStream_t s) {
  if (B.n < 1) return;
  if (B.n > EIG_BATCH_MAX) throw std::invalid_argument("eig_cg_batch: too many levels");
  if (weighted) {
    if (ncomp != 2) throw std::invalid_argument("weighted mass operator expects 2 components");
    hipLaunchKernelGGL((k_eig_cg<K, 2, true>), dim3(B.n), dim3(1024), 0, s, B, coeff, nit, tol);
  } else if (ncomp == 1) {
    hipLaunchKernelGGL((k_eig_cg<K, 1, false>), dim3(B.n), dim3(1024), 0, s, B, coeff, nit, tol);
  } else {
    hipLaunchKernelGGL((k_eig_cg<K, 2, false>), dim3(B.n), dim3(1024), 0, s, B, coeff, nit, tol);
  }
  ATMO_HIP_CHECK(hipGetLastError());
}

#define ATMO_MG_INST(K)                                                                                          \
  template void launch_restrict<K, double>(const MeshDev&, const MeshDev&, const double*, double*, int, long, long, \
                                           hipStream_t);                                                        \
  template void launch_restrict<K, float>(const MeshDev&, const MeshDev&, const float*, float*, int, long, long,   \
                                          hipStream_t);                                                         \
  template void launch_project_down<K>(const MeshDev&, const MeshDev&, const double*, double*, hipStream_t);      \
  template void launch_presmooth_residual<K, float>(const MeshDev&, int, bool, float*, const float*, float,          \
                                                    const float*, const float*, const ChebSteps&, float*, long,      \
                                                    bool, hipStream_t);                                                    \
  template void launch_presmooth_residual<K, double>(const MeshDev&, int, bool, double*, const double*, double,      \
                                                     const double*, const double*, const ChebSteps&, double*, long,  \
                                                     bool, hipStream_t);                                                   \
  template void launch_prolongate_postsmooth<K, float>(const MeshDev&, const MeshDev&, int, bool, float*,            \
                                                       const float*, const float*, float, const float*, const float*, \
                                                       const ChebSteps&, long, long, bool, hipStream_t);             \
  template void launch_prolongate_postsmooth<K, double>(const MeshDev&, const MeshDev&, int, bool, double*,          \
                                                        const double*, const double*, double, const double*,         \
                                                        const double*, const ChebSteps&, long, long, bool,           \
                                                        hipStream_t);                                                \
  template void launch_coarse_cg_f32<K>(const MeshDev&, int, bool, const double*, double, const float*, float*,   \
                                        double*, double*, double*, double*, double*, const double*, int, int*, long, \
                                        hipStream_t);                                                              \
  template void launch_coarse_cg<K>(const MeshDev&, int, bool, const double*, double, const double*, double*,     \
                                    double*, double*, double*, const double*, int, int*, long, hipStream_t);       \
  template void launch_eig_cg_batch<K>(const EigBatch&, int, bool, double, int, double, hipStream_t);

ATMO_MG_INST(1)
ATMO_MG_INST(2)
ATMO_MG_INST(3)


}  // namespace atmo
