// Operator kernels of the IMEX-DG hot path (euler_operator.h).
//
// Two kernel families:
//  * k_mass / k_inv_diag: the cell-local mass-type blocks M_rho, D and A
//    (euler_operator.h:717-740, 1384-1408, 2328-2357) and their diagonals
//    (:2456-2626), templated on the arithmetic type (FP64 and FP32 levels).
//  * k_cellface<Phys>: cell integral plus the cell's own four faces for the
//    operators with face terms (B, C and the explicit residuals). Each cell
//    evaluates the flux of each of its faces with itself as the "+" side and
//    its outward normal. All the reference's face fluxes are antisymmetric
//    under exchanging the sides (n -> -n), so this equals the reference's
//    face-loop contribution to the cell without any atomics; every face flux
//    is computed twice (once per side).
#include "cellface.h"
#include "ops.h"

namespace atmo {

Tableau::Tableau() {
  g = 2.0 - std::sqrt(2.0);
  a21 = g;
  a22 = 0.0;
  a31 = 0.5;
  a32 = 0.5;
  a33 = 0.0;
  a21t = 0.5 * g;
  a22t = 0.5 * g;
  a31t = std::sqrt(2.0) / 4.0;
  a32t = std::sqrt(2.0) / 4.0;
  a33t = 1.0 - std::sqrt(2.0) / 2.0;
  b1 = a31t;
  b2 = a32t;
  b3 = a33t;
}

// ------------------------------------------------------------ mass family
template <int K, typename T, int NC, bool W>
__global__ __launch_bounds__(64) void k_mass(MeshDev m, const T* __restrict__ src, const T* __restrict__ rho,
                                             T* __restrict__ dst, T coeff, const T* __restrict__ add, T add_scale,
                                             long Ns) {
  using C = Cfg<K>;
  constexpr int ND2 = C::ND2, CPW = C::CPW;
  constexpr int NS = NC + (W ? 1 : 0);
  __shared__ T s_u[NS][CPW * ND2];  // the components, then rho when W
  __shared__ T s_red[NC][CPW][C::ND][C::NQ];
  const T* srcs[NS];
#pragma unroll
  for (int comp = 0; comp < NC; ++comp) srcs[comp] = src + comp * Ns;
  if constexpr (W) srcs[NS - 1] = rho;
  const MassLane<K, T> pre = wave_mass_lane<K, T>(m, cell_block() * CPW);
  stage_fields<K, NS, T, T>(&s_u[0][0], srcs, m.ncell);
  __syncthreads();
  wave_mass_core<K, T, NC, W>(m, cell_block() * CPW, &s_u[0][0], &s_u[NS - 1][0], coeff, &s_red[0][0][0][0], &pre);
  __syncthreads();
  const long base = (long)cell_block() * CPW * ND2;
  const long lim = (long)m.ncell * ND2;
  T Si[C::NQ];
  lane_cols_S<K, T>(threadIdx.x % C::ND, Si);
  for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
    const long gi = base + t;
    if (gi >= lim) break;
    if (64 % C::ND != 0) lane_cols_S<K, T>(t % C::ND, Si);
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) {
      T acc = wave_mass_out<K, T>(&s_red[comp][0][0][0], t, Si);
      if (add) acc += add_scale * add[comp * Ns + gi];
      dst[comp * Ns + gi] = acc;
    }
  }
}

template <int K, typename T, bool W>
__global__ __launch_bounds__(64) void k_inv_diag(MeshDev m, int ncomp, const T* __restrict__ rho, T* __restrict__ dst,
                                                 T coeff, long Ns) {
  using C = Cfg<K>;
  constexpr int ND = C::ND, NQ = C::NQ, ND2 = C::ND2, CPW = C::CPW;
  constexpr auto& tb = basis::kT<K>;
  __shared__ T s_r[W ? CPW * ND2 : 1];
  __shared__ T s_red[CPW][ND][NQ];
  LaneMap<K> L(m.ncell);
  if constexpr (W) {
    const T* s1[1] = {rho};
    stage_fields<K, 1, T, T>(s_r, s1, m.ncell);
  }
  __syncthreads();
  if (L.active) {
    const CellVerts v = cell_verts(m, L.cell);
    const ColGeo<T> g = col_geo<K, T>(v, L.c);
    T Sc[ND], Dc[ND];
    lane_rows<K, T>(L.c, Sc, Dc);
    T wq[NQ];
    const T jc = g.dx * g.ze * g.wc * coeff;
#pragma unroll
    for (int r = 0; r < NQ; ++r) wq[r] = jc * (T)tb.gw[r];
    if (W) {
      T tv[ND];
#pragma unroll
      for (int j = 0; j < ND; ++j) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < ND; ++i) s += Sc[i] * s_r[L.cw * ND2 + j * ND + i];
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
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      T s = 0;
#pragma unroll
      for (int r = 0; r < NQ; ++r) s += (T)(tb.S[r][j] * tb.S[r][j]) * wq[r];
      s_red[L.cw][j][L.c] = s;
    }
  }
  __syncthreads();
  const long base = (long)cell_block() * CPW * ND2;
  const long lim = (long)m.ncell * ND2;
  for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
    const long gi = base + t;
    if (gi >= lim) break;
    const int cw = t / ND2, o = t - cw * ND2, jj = o / ND, ii = o - jj * ND;
    T Si[NQ];
    lane_cols_S<K, T>(ii, Si);
    T acc = 0;
#pragma unroll
    for (int c = 0; c < NQ; ++c) acc += Si[c] * Si[c] * s_red[cw][jj][c];
    const T inv = (T)1 / acc;
    for (int comp = 0; comp < ncomp; ++comp) dst[comp * Ns + gi] = inv;
  }
}

template <int K, typename T>
void launch_mass(const MeshDev& m, int ncomp, bool weighted, const T* src, const T* rho, T* dst, T coeff,
                 const T* add, T add_scale, long Ns, hipStream_t s) {
  const int grid = ceil_div(m.ncell, Cfg<K>::CPW);
  if (grid == 0) return;
  if (weighted) {
    if (ncomp != 2) throw std::invalid_argument("weighted mass operator expects 2 components");
    hipLaunchKernelGGL((k_mass<K, T, 2, true>), dim3(grid), dim3(64), 0, s, m, src, rho, dst, coeff, add,
                       add_scale, Ns);
  } else if (ncomp == 1) {
    hipLaunchKernelGGL((k_mass<K, T, 1, false>), dim3(grid), dim3(64), 0, s, m, src, rho, dst, coeff, add,
                       add_scale, Ns);
  } else {
    hipLaunchKernelGGL((k_mass<K, T, 2, false>), dim3(grid), dim3(64), 0, s, m, src, rho, dst, coeff, add,
                       add_scale, Ns);
  }
  ATMO_HIP_CHECK(hipGetLastError());
}

template <int K, typename T>
void launch_inv_diag(const MeshDev& m, int ncomp, bool weighted, const T* rho, T* dst, T coeff, long Ns,
                     hipStream_t s) {
  const int grid = ceil_div(m.ncell, Cfg<K>::CPW);
  if (grid == 0) return;
  if (weighted)
    hipLaunchKernelGGL((k_inv_diag<K, T, true>), dim3(grid), dim3(64), 0, s, m, ncomp, rho, dst, coeff, Ns);
  else
    hipLaunchKernelGGL((k_inv_diag<K, T, false>), dim3(grid), dim3(64), 0, s, m, ncomp, rho, dst, coeff, Ns);
  ATMO_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------- cell+face family
// density residual: euler_operator.h:344-692
template <int ST>
struct PhysRhsRho {
  static constexpr int NF = 3 * ST, NO = 1;  // (rho, ux, uz) per stage
  static constexpr bool BDRY = false;
  static constexpr bool HAS_V = true;
  static constexpr unsigned GX = (1u << NO) - 1, GZ = (1u << NO) - 1;
  double wdt[3];
  __device__ void cell(const double (&v)[NF], double (&V)[NO], double (&Gx)[NO], double (&Gz)[NO]) const {
    double gx = 0, gz = 0;
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      gx += wdt[s] * v[3 * s] * v[3 * s + 1];
      gz += wdt[s] * v[3 * s] * v[3 * s + 2];
    }
    V[0] = v[0];
    Gx[0] = gx;
    Gz[0] = gz;
  }
  __device__ void face(const double (&a)[NF], const double (&b)[NF], double n0, double n1, double (&F)[NO]) const {
    double acc = 0;
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      const double ra = a[3 * s], rb = b[3 * s];
      const double ax = 0.5 * (ra * a[3 * s + 1] + rb * b[3 * s + 1]);
      const double az = 0.5 * (ra * a[3 * s + 2] + rb * b[3 * s + 2]);
      const double lam = fmax(fabs(a[3 * s + 1] * n0 + a[3 * s + 2] * n1), fabs(b[3 * s + 1] * n0 + b[3 * s + 2] * n1));
      acc += wdt[s] * (ax * n0 + az * n1 + 0.5 * lam * (ra - rb));
    }
    F[0] = -acc;
  }
  __device__ void bdry(const double (&)[NF], double, double, double (&F)[NO]) const { F[0] = 0; }
};

// energy (pressure) residual: euler_operator.h:745-1361
template <int ST>
struct PhysRhsP {
  static constexpr int NF = 4 * ST + 3, NO = 1;  // (rho,ux,uz,p) per stage, then rho_k, uk_x, uk_z
  static constexpr bool BDRY = false;
  static constexpr bool HAS_V = true;
  static constexpr unsigned GX = (1u << NO) - 1, GZ = (1u << NO) - 1;
  double wa[3], wt[3], gwt[3];  // advective, enthalpy and gravity weights
  double dt, Ma2, gm, iFr2, G, e0, e1;
  __device__ void cell(const double (&v)[NF], double (&V)[NO], double (&Gx)[NO], double (&Gz)[NO]) const {
    constexpr int KB = 4 * ST;
    const double rk = v[KB], ukx = v[KB + 1], ukz = v[KB + 2];
    double E[ST], uu[ST];
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      uu[s] = v[4 * s + 1] * v[4 * s + 1] + v[4 * s + 2] * v[4 * s + 2];
      E[s] = 1.0 / (gm - 1.0) * (v[4 * s + 3] / v[4 * s]) + 0.5 * Ma2 * uu[s];
    }
    double val = v[0] * E[0] - 0.5 * rk * Ma2 * (ukx * ukx + ukz * ukz);
    // gravity work: explicit stages use their own (rho, u); the implicit
    // weight uses (rho_k, u_k) in stages 1 and 2
#pragma unroll
    for (int s = 0; s < ST; ++s)
      if (ST == 3 || s < ST)
        val -= gwt[s] * dt * Ma2 * iFr2 * v[4 * s] * (v[4 * s + 1] * e0 + v[4 * s + 2] * e1) * G;
    if (ST < 3) val -= gwt[ST] * dt * Ma2 * iFr2 * rk * (ukx * e0 + ukz * e1) * G;
    double gx = 0, gz = 0;
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      const double kin = 0.5 * wa[s] * dt * Ma2 * uu[s] * v[4 * s];
      const double ent = wt[s] * dt * (v[4 * s] * (E[s] - 0.5 * Ma2 * uu[s]) + v[4 * s + 3]);
      gx += kin * v[4 * s + 1] + ent * v[4 * s + 1];
      gz += kin * v[4 * s + 2] + ent * v[4 * s + 2];
    }
    V[0] = val;
    Gx[0] = gx;
    Gz[0] = gz;
  }
  __device__ void face(const double (&a)[NF], const double (&b)[NF], double n0, double n1, double (&F)[NO]) const {
    double acc = 0;
#pragma unroll
    for (int s = 0; s < ST; ++s) {
      const double ra = a[4 * s], rb = b[4 * s];
      const double uxa = a[4 * s + 1], uza = a[4 * s + 2], uxb = b[4 * s + 1], uzb = b[4 * s + 2];
      const double pa = a[4 * s + 3], pb = b[4 * s + 3];
      const double ua2 = uxa * uxa + uza * uza, ub2 = uxb * uxb + uzb * uzb;
      const double akx = 0.5 * (0.5 * ua2 * ra * uxa + 0.5 * ub2 * rb * uxb);
      const double akz = 0.5 * (0.5 * ua2 * ra * uza + 0.5 * ub2 * rb * uzb);
      const double Ea = 1.0 / (gm - 1.0) * pa / ra + 0.5 * Ma2 * ua2;
      const double Eb = 1.0 / (gm - 1.0) * pb / rb + 0.5 * Ma2 * ub2;
      const double ha = (Ea - 0.5 * Ma2 * ua2) * ra + pa;
      const double hb = (Eb - 0.5 * Ma2 * ub2) * rb + pb;
      const double aex = 0.5 * (ha * uxa + hb * uxb);
      const double aez = 0.5 * (ha * uza + hb * uzb);
      const double lam = fmax(fabs(uxa * n0 + uza * n1), fabs(uxb * n0 + uzb * n1));
      if (ST < 3) {
        double jkin;
        if (ST == 2 && s == 1) {
          // euler_operator.h:1182-1183: u_tmp_2 . u_old, as written
          jkin = ra * 0.5 * (uxa * a[1] + uza * a[2]) - rb * 0.5 * (uxb * b[1] + uzb * b[2]);
        } else {
          jkin = ra * 0.5 * ua2 - rb * 0.5 * ub2;
        }
        const double je = ra * (Ea - 0.5 * Ma2 * ua2) - rb * (Eb - 0.5 * Ma2 * ub2);
        acc += -wa[s] * dt * Ma2 * (akx * n0 + akz * n1 + 0.5 * lam * jkin) - wt[s] * dt * (aex * n0 + aez * n1 + 0.5 * lam * je);
      } else {
        const double jrE = ra * Ea - rb * Eb;
        acc += -wa[s] * dt * Ma2 * (akx * n0 + akz * n1) - wa[s] * dt * (aex * n0 + aez * n1) - wa[s] * dt * 0.5 * lam * jrE;
      }
    }
    F[0] = acc;
  }
  __device__ void bdry(const double (&)[NF], double, double, double (&F)[NO]) const { F[0] = 0; }
};

// momentum (velocity) residual: euler_operator.h:1620-2305
template <int ST>
struct PhysRhsU {
  static constexpr int NX = ST < 3 ? 1 : 0;  // extra gravity density (rho_2 resp. rho_3)
  static constexpr int NF = 4 * ST + NX, NO = 2;
  static constexpr bool BDRY = true;
  static constexpr bool HAS_V = true;
  static constexpr unsigned GX = (1u << NO) - 1, GZ = (1u << NO) - 1;
  double wa[3], wt[3], gwt[3];
  double dt, Ma2, iFr2, G, e0, e1;
  __device__ void cell(const double (&v)[NF], double (&V)[NO], double (&Gx)[NO], double (&Gz)[NO]) const {
    double grho = 0;  // sum of gravity weight * density
#pragma unroll
    for (int s = 0; s < ST; ++s) grho += gwt[s] * v[4 * s];
    if (NX) grho += gwt[ST] * v[4 * ST];
    (void)grho;
    double gs[NO];
#pragma unroll
    for (int comp = 0; comp < NO; ++comp) {
      const double e = comp == 0 ? e0 : e1;
      double val = v[0] * v[1 + comp];
#pragma unroll
      for (int s = 0; s < ST; ++s) val -= gwt[s] * dt * iFr2 * v[4 * s] * e * G;
      if (NX) val -= gwt[ST] * dt * iFr2 * v[4 * ST] * e * G;
      V[comp] = val;
      double gx = 0, gz = 0;
#pragma unroll
      for (int s = 0; s < ST; ++s) {
        const double uc = v[4 * s + 1 + comp];
        gx += wa[s] * dt * v[4 * s] * uc * v[4 * s + 1] + (comp == 0 ? wt[s] * dt / Ma2 * v[4 * s + 3] : 0.0);
        gz += wa[s] * dt * v[4 * s] * uc * v[4 * s + 2] + (comp == 1 ? wt[s] * dt / Ma2 * v[4 * s + 3] : 0.0);
      }
      Gx[comp] = gx;
      Gz[comp] = gz;
      gs[comp] = 0;
    }
    (void)gs;
  }
  __device__ void face(const double (&a)[NF], const double (&b)[NF], double n0, double n1, double (&F)[NO]) const {
#pragma unroll
    for (int comp = 0; comp < NO; ++comp) {
      const double nc = comp == 0 ? n0 : n1;
      double acc = 0;
#pragma unroll
      for (int s = 0; s < ST; ++s) {
        const double ra = a[4 * s], rb = b[4 * s];
        const double uxa = a[4 * s + 1], uza = a[4 * s + 2], uxb = b[4 * s + 1], uzb = b[4 * s + 2];
        const double uca = a[4 * s + 1 + comp], ucb = b[4 * s + 1 + comp];
        const double tn = 0.5 * (ra * uca * (uxa * n0 + uza * n1) + rb * ucb * (uxb * n0 + uzb * n1));
        const double avgp = 0.5 * (a[4 * s + 3] + b[4 * s + 3]);
        const double jr = ra * uca - rb * ucb;
        const double lam = fmax(fabs(uxa * n0 + uza * n1), fabs(uxb * n0 + uzb * n1));
        acc += -wa[s] * dt * tn - wt[s] * dt / Ma2 * avgp * nc - wa[s] * dt * 0.5 * lam * jr;
      }
      F[comp] = acc;
    }
  }
  __device__ void bdry(const double (&a)[NF], double n0, double n1, double (&F)[NO]) const {
#pragma unroll
    for (int comp = 0; comp < NO; ++comp) {
      const double nc = comp == 0 ? n0 : n1;
      double acc = 0;
#pragma unroll
      for (int s = 0; s < ST; ++s) {
        const double pD = a[4 * s + 3];
        acc += -wt[s] * dt / Ma2 * 0.5 * (a[4 * s + 3] + pD) * nc;
      }
      F[comp] = acc;
    }
  }
};

template <int K, class Ph, bool AMR>
__global__ __launch_bounds__(64) void k_cellface(MeshDev m, Ph ph, FieldPtrs<Ph::NF> in, double* __restrict__ dst,
                                                 const double* __restrict__ scale, long Ns) {
  constexpr int ND2 = Cfg<K>::ND2, CPW = Cfg<K>::CPW;
  __shared__ CFSmem<K, Ph, AMR> sh;
  cellface_columns<K, Ph, AMR>(m, ph, in, sh);
  const long base = (long)cell_block() * CPW * ND2;
  const long lim = (long)m.ncell * ND2;
  OutRows<K> R = out_rows<K>(threadIdx.x);
  for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
    const long gi = base + t;
    if (gi >= lim) break;
    if (!rows_invariant<K>()) R = out_rows<K>(t);
#pragma unroll
    for (int o = 0; o < Ph::NO; ++o) {
      double acc = cellface_out<K, Ph, AMR>(sh, o, t, R);
      if (scale) acc *= scale[o * Ns + gi];
      dst[o * Ns + gi] = acc;
    }
  }
}

template <int K, class Ph>
static void launch_cf(const MeshDev& m, const Ph& ph, const FieldPtrs<Ph::NF>& in, double* dst, long Ns,
                      hipStream_t s, const double* scale = nullptr) {
  const int grid = ceil_div(m.ncell, Cfg<K>::CPW);
  if (grid == 0) return;
  if (m.cface)
    hipLaunchKernelGGL((k_cellface<K, Ph, true>), dim3(grid), dim3(64), 0, s, m, ph, in, dst, scale, Ns);
  else
    hipLaunchKernelGGL((k_cellface<K, Ph, false>), dim3(grid), dim3(64), 0, s, m, ph, in, dst, scale, Ns);
  ATMO_HIP_CHECK(hipGetLastError());
}

template <int K>
void launch_B(const MeshDev& m, const OpConsts& oc, double coeff, const double* p, double* dst_u, long Ns,
              hipStream_t s) {
  PhysB ph;
  ph.cf = coeff * oc.dt / (oc.Ma * oc.Ma);
  FieldPtrs<1> in{{p}};
  launch_cf<K>(m, ph, in, dst_u, Ns, s);
}

template <int K>
void launch_C(const MeshDev& m, const OpConsts& oc, double coeff, const double* pfix, const double* u,
              double* dst_p, long Ns, hipStream_t s) {
  PhysC ph;
  ph.set(coeff * oc.dt, oc.gamma_gas);
  FieldPtrs<3> in{{pfix, u, u + Ns}};
  launch_cf<K>(m, ph, in, dst_p, Ns, s);
}

template <int K>
void launch_d_minus_c(const MeshDev& m, const OpConsts& oc, double coeff, const double* pfix, const double* u,
                      const double* p, const double* scale, double* dst, long Ns, hipStream_t s) {
  PhysDminusC ph;
  ph.c.set(coeff * oc.dt, oc.gamma_gas);
  ph.cD = 1.0 / (oc.gamma_gas - 1.0);
  FieldPtrs<4> in{{pfix, u, u + Ns, p}};
  launch_cf<K>(m, ph, in, dst, Ns, s, scale);
}

// w = scale * (D p - C u) (as k_cellface<PhysDminusC>) and, in the same pass,
// the partial dot products of w with V_0 .. V_{nv-1} (column k at V + k*ldv)
// and w.w: slot k of block b at part[k*pstride + b], slot nv is w.w. The
// Arnoldi step of GMRES: its Schur product and the classical Gram-Schmidt dot
// products without a second pass over w. With p_last, p is column nv-1 of V
// (GMRES applies the operator to its newest basis vector): that slot's values
// come from the staged copy of p instead of a second read.
// WPE: the waves per SIMD the kernel is compiled for (its register budget), a tuning choice with identical
// results (arnoldi_waves_per_simd; 0 picks the measured default: 3 at k = 2, where 4 spills, else 4).
template <int K, bool AMR, int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(AMR ? 1 : WPE))) void k_dmc_dots(MeshDev m, PhysDminusC ph, FieldPtrs<4> in, double* __restrict__ dst,
                                                 const double* __restrict__ scale, long Ns,
                                                 const double* __restrict__ V, long ldv, int nv, int p_last,
                                                 double* __restrict__ part, long pstride) {
  constexpr int ND2 = Cfg<K>::ND2, CPW = Cfg<K>::CPW, NPASS = (CPW * ND2 + 63) / 64;
  using SM = CFSmem<K, PhysDminusC, AMR>;
  __shared__ SM sh;
  // the block's own p values (field 3 of the staged inputs), kept past the column phase
  __shared__ double pk[CPW * ND2];
  const int blk = cell_block();
  const long base = (long)blk * CPW * ND2;
  const long lim = (long)m.ncell * ND2;
  cellface_columns<K, PhysDminusC, AMR>(m, ph, in, sh, NoPrefetch(), [&]() {
    if (!p_last) return;
#pragma unroll
    for (int ps = 0; ps < NPASS; ++ps) {
      const int t = threadIdx.x + 64 * ps;
      if (t < CPW * ND2) pk[t] = sh.in[3][(t / ND2 + SM::HALO) * SM::CS + t % ND2];
    }
  });
  // the Jacobi scale of the block's dofs, loaded once the column phase is done (loaded with the staged
  // fields, it stays live through the column phase, whose registers then spill: measured slower)
  double sc[NPASS];
#pragma unroll
  for (int ps = 0; ps < NPASS; ++ps) {
    const int t = threadIdx.x + 64 * ps;
    sc[ps] = (t < CPW * ND2 && base + t < lim) ? scale[base + t] : 0.0;
  }
  OutRows<K> R = out_rows<K>(threadIdx.x);
  // the block's dofs: full passes of 64 lanes, then a tail of TL < 64 dofs
  constexpr int NFULL = (CPW * ND2) / 64, TL = CPW * ND2 - 64 * NFULL;
  double wv[NPASS];
  long gi[NPASS];
#pragma unroll
  for (int ps = 0; ps < NPASS; ++ps) {
    const int t = threadIdx.x + 64 * ps;
    gi[ps] = base + t;
    wv[ps] = 0.0;
    if (t < CPW * ND2 && gi[ps] < lim) {
      if (!rows_invariant<K>()) R = out_rows<K>(t);
      const double acc = cellface_out<K, PhysDminusC, AMR>(sh, 0, t, R) * sc[ps];
      dst[gi[ps]] = acc;
      wv[ps] = acc;
    } else {
      gi[ps] = base;  // in bounds (the block's first dof); its product is zero
    }
  }
  const int l = threadIdx.x & 63;
  // Dot products. Slots 0 .. nload-1 are basis columns read from memory, in chunks of 8 whose loads are
  // all in flight at once (slots past nload load nothing and multiply zeros). Full passes: every lane
  // multiplies its own dofs. Tail (TL dofs, TL/8 per slot when TL divides into 8 lanes' pairs): lanes
  // 8s .. 8s+7 take slot s's tail as one 16-byte load of two dofs, with w's tail values fetched from the
  // lanes that formed them, so a chunk takes 8*NFULL + 1 loads per lane instead of 8*NPASS. The basis is
  // read with non-temporal loads (streamed, not kept in the caches). The other two slots need no loads
  // and are reduced together after the chunks: with p_last, slot nv-1 is p (the staged copy), and slot
  // nv is w itself.
  constexpr bool PAIRS = TL == 16;
  constexpr int NL = PAIRS ? NFULL : NPASS;  // per-lane single loads per slot
  double wt0 = 0.0, wt1 = 0.0;
  long gt = base;
  if constexpr (PAIRS) {
    const int src = 2 * (l & 7);
    wt0 = __shfl(wv[NFULL], src, 64);
    wt1 = __shfl(wv[NFULL], src + 1, 64);
    gt = base + 64 * NFULL + src;
    if (gt + 1 >= lim) gt = base;  // past the last cell: w's values there are zero
  }
  const int nload = p_last ? nv - 1 : nv;  // the slots read from memory
  typedef double nt_d2 __attribute__((ext_vector_type(2)));
  struct Chunk {
    double y[8][NL];
    nt_d2 t;
  };
  auto load = [&](int k0, Chunk& c) {
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      if (k0 + s < nload) {  // uniform
        const long col = (long)(k0 + s) * ldv;
#pragma unroll
        for (int ps = 0; ps < NL; ++ps) c.y[s][ps] = __builtin_nontemporal_load(V + col + gi[ps]);
      } else {
#pragma unroll
        for (int ps = 0; ps < NL; ++ps) c.y[s][ps] = 0.0;
      }
    }
    c.t = nt_d2{0.0, 0.0};
    if constexpr (PAIRS) {
      const int k = k0 + (l >> 3);
      if (k < nload) c.t = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(V + (long)k * ldv + gt));
    }
  };
  auto reduce = [&](int k0, const Chunk& c) {
    double v[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      double acc = 0.0;
#pragma unroll
      for (int ps = 0; ps < NL; ++ps) acc += wv[ps] * c.y[s][ps];
      v[s] = acc;
    }
    if constexpr (PAIRS) {
      const double tail = wt0 * c.t.x + wt1 * c.t.y;
#pragma unroll
      for (int s = 0; s < 8; ++s)
        if (s == (l >> 3)) v[s] += tail;
    }
    const double sum = wave_sum8(v);
    const int slot = k0 + (l >> 3);
    if ((l & 7) == 0 && slot < nload) part[(long)slot * pstride + blk] = sum;
  };
  // w.p (with p_last) and w.w: this lane's parts from LDS and registers, reduced together while the first
  // chunk's loads are in flight
  auto specials = [&]() {
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ps = 0; ps < NL; ++ps) {
      const int t = threadIdx.x + 64 * ps;  // past the block's dofs w is zero and pk has no entry
      if (p_last && t < CPW * ND2) v[0] += wv[ps] * pk[t];
      v[1] += wv[ps] * wv[ps];
    }
    if constexpr (PAIRS) {
      if (l < 8) {
        const int tl = 64 * NFULL + 2 * l;  // the block-local dofs of this lane's tail pair
        if (p_last) v[0] += wt0 * pk[tl] + wt1 * pk[tl + 1];
        v[1] += wt0 * wt0 + wt1 * wt1;
      }
    }
    const double sum = wave_sum8(v);
    if (l == 0 && p_last) part[(long)(nv - 1) * pstride + blk] = sum;
    if (l == 8) part[(long)nv * pstride + blk] = sum;
  };
  if (nload == 0) {
    specials();
    return;
  }
  for (int k0 = 0; k0 < nload; k0 += 8) {
    Chunk c;
    load(k0, c);
    if (k0 == 0) specials();
    reduce(k0, c);
  }
}

template <int K>
void launch_d_minus_c_dots(const MeshDev& m, const OpConsts& oc, double coeff, const double* pfix, const double* u,
                           const double* p, const double* scale, double* dst, long Ns, const double* V, long ldv,
                           int nv, double* part, long pstride, int waves, hipStream_t s) {
  // GMRES applies the operator to its newest basis vector: then slot nv-1 is p itself
  const int p_last = nv > 0 && p == V + (long)(nv - 1) * ldv;
  const int grid = ceil_div(m.ncell, Cfg<K>::CPW);
  if (grid == 0) return;
  if (grid > pstride) throw std::invalid_argument("d_minus_c_dots: partial buffer too small");
  PhysDminusC ph;
  ph.c.set(coeff * oc.dt, oc.gamma_gas);
  ph.cD = 1.0 / (oc.gamma_gas - 1.0);
  FieldPtrs<4> in{{pfix, u, u + Ns, p}};
  if (waves == 0) waves = K == 2 ? 3 : 4;
  if (waves != 3 && waves != 4) throw std::invalid_argument("d_minus_c_dots: waves per SIMD must be 0, 3 or 4");
  if (m.cface)
    hipLaunchKernelGGL((k_dmc_dots<K, true, 1>), dim3(grid), dim3(64), 0, s, m, ph, in, dst, scale, Ns, V, ldv, nv,
                       p_last, part, pstride);
  else if (waves == 3)
    hipLaunchKernelGGL((k_dmc_dots<K, false, 3>), dim3(grid), dim3(64), 0, s, m, ph, in, dst, scale, Ns, V, ldv,
                       nv, p_last, part, pstride);
  else
    hipLaunchKernelGGL((k_dmc_dots<K, false, 4>), dim3(grid), dim3(64), 0, s, m, ph, in, dst, scale, Ns, V, ldv,
                       nv, p_last, part, pstride);
  ATMO_HIP_CHECK(hipGetLastError());
}

template <int K>
void launch_rhs_rho(const MeshDev& m, const OpConsts& oc, const Tableau& tb, int stage, const double* const* src,
                    double* dst, long Ns, hipStream_t s) {
  if (stage == 1) {
    PhysRhsRho<1> ph;
    ph.wdt[0] = tb.a21 * oc.dt;
    FieldPtrs<3> in{{src[0], src[1], src[1] + Ns}};
    launch_cf<K>(m, ph, in, dst, Ns, s);
  } else if (stage == 2) {
    PhysRhsRho<2> ph;
    ph.wdt[0] = tb.a31 * oc.dt;
    ph.wdt[1] = tb.a32 * oc.dt;
    FieldPtrs<6> in{{src[0], src[1], src[1] + Ns, src[2], src[3], src[3] + Ns}};
    launch_cf<K>(m, ph, in, dst, Ns, s);
  } else if (stage == 3) {
    PhysRhsRho<3> ph;
    ph.wdt[0] = tb.b1 * oc.dt;
    ph.wdt[1] = tb.b2 * oc.dt;
    ph.wdt[2] = tb.b3 * oc.dt;
    FieldPtrs<9> in{{src[0], src[1], src[1] + Ns, src[2], src[3], src[3] + Ns, src[4], src[5], src[5] + Ns}};
    launch_cf<K>(m, ph, in, dst, Ns, s);
  } else {
    throw std::invalid_argument("rhs_rho: stage must be 1..3");
  }
}

template <int ST>
static void fill_weights(const Tableau& tb, double (&wa)[3], double (&wt)[3], double (&gw)[3]) {
  if (ST == 1) {
    wa[0] = tb.a21; wt[0] = tb.a21t;
    gw[0] = tb.a21t; gw[1] = tb.a22t;
  } else if (ST == 2) {
    wa[0] = tb.a31; wt[0] = tb.a31t; wa[1] = tb.a32; wt[1] = tb.a32t;
    gw[0] = tb.a31t; gw[1] = tb.a32t; gw[2] = tb.a33t;
  } else {
    wa[0] = wt[0] = tb.b1; wa[1] = wt[1] = tb.b2; wa[2] = wt[2] = tb.b3;
    gw[0] = tb.b1; gw[1] = tb.b2; gw[2] = tb.b3;
  }
}

template <int K, int ST>
static void rhs_p_stage(const MeshDev& m, const OpConsts& oc, const Tableau& tb, const double* const* src,
                        double* dst, long Ns, hipStream_t s) {
  PhysRhsP<ST> ph;
  fill_weights<ST>(tb, ph.wa, ph.wt, ph.gwt);
  ph.dt = oc.dt;
  ph.Ma2 = oc.Ma * oc.Ma;
  ph.gm = oc.gamma_gas;
  ph.iFr2 = 1.0 / (oc.Fr * oc.Fr);
  ph.G = oc.grav_factor;
  ph.e0 = oc.grav[0];
  ph.e1 = oc.grav[1];
  FieldPtrs<PhysRhsP<ST>::NF> in;
  for (int t = 0; t < ST; ++t) {
    in.p[4 * t] = src[3 * t];
    in.p[4 * t + 1] = src[3 * t + 1];
    in.p[4 * t + 2] = src[3 * t + 1] + Ns;
    in.p[4 * t + 3] = src[3 * t + 2];
  }
  in.p[4 * ST] = src[3 * ST];
  in.p[4 * ST + 1] = src[3 * ST + 1];
  in.p[4 * ST + 2] = src[3 * ST + 1] + Ns;
  launch_cf<K>(m, ph, in, dst, Ns, s);
}

template <int K>
void launch_rhs_pressure(const MeshDev& m, const OpConsts& oc, const Tableau& tb, int stage,
                         const double* const* src, double* dst, long Ns, hipStream_t s) {
  if (stage == 1) rhs_p_stage<K, 1>(m, oc, tb, src, dst, Ns, s);
  else if (stage == 2) rhs_p_stage<K, 2>(m, oc, tb, src, dst, Ns, s);
  else if (stage == 3) rhs_p_stage<K, 3>(m, oc, tb, src, dst, Ns, s);
  else throw std::invalid_argument("rhs_pressure: stage must be 1..3");
}

template <int K, int ST>
static void rhs_u_stage(const MeshDev& m, const OpConsts& oc, const Tableau& tb, const double* const* src,
                        double* dst, long Ns, hipStream_t s) {
  PhysRhsU<ST> ph;
  fill_weights<ST>(tb, ph.wa, ph.wt, ph.gwt);
  ph.dt = oc.dt;
  ph.Ma2 = oc.Ma * oc.Ma;
  ph.iFr2 = 1.0 / (oc.Fr * oc.Fr);
  ph.G = oc.grav_factor;
  ph.e0 = oc.grav[0];
  ph.e1 = oc.grav[1];
  FieldPtrs<PhysRhsU<ST>::NF> in;
  for (int t = 0; t < ST; ++t) {
    in.p[4 * t] = src[3 * t];
    in.p[4 * t + 1] = src[3 * t + 1];
    in.p[4 * t + 2] = src[3 * t + 1] + Ns;
    in.p[4 * t + 3] = src[3 * t + 2];
  }
  if (PhysRhsU<ST>::NX) in.p[4 * ST] = src[3 * ST];
  launch_cf<K>(m, ph, in, dst, Ns, s);
}

template <int K>
void launch_rhs_velocity(const MeshDev& m, const OpConsts& oc, const Tableau& tb, int stage,
                         const double* const* src, double* dst, long Ns, hipStream_t s) {
  if (stage == 1) rhs_u_stage<K, 1>(m, oc, tb, src, dst, Ns, s);
  else if (stage == 2) rhs_u_stage<K, 2>(m, oc, tb, src, dst, Ns, s);
  else if (stage == 3) rhs_u_stage<K, 3>(m, oc, tb, src, dst, Ns, s);
  else throw std::invalid_argument("rhs_velocity: stage must be 1..3");
}

#define ATMO_INST_K(K)                                                                                          \
  template void launch_mass<K, double>(const MeshDev&, int, bool, const double*, const double*, double*, double, \
                                       const double*, double, long, hipStream_t);                               \
  template void launch_mass<K, float>(const MeshDev&, int, bool, const float*, const float*, float*, float,      \
                                      const float*, float, long, hipStream_t);                                  \
  template void launch_inv_diag<K, double>(const MeshDev&, int, bool, const double*, double*, double, long,      \
                                           hipStream_t);                                                        \
  template void launch_inv_diag<K, float>(const MeshDev&, int, bool, const float*, float*, float, long,          \
                                          hipStream_t);                                                         \
  template void launch_B<K>(const MeshDev&, const OpConsts&, double, const double*, double*, long, hipStream_t);   \
  template void launch_C<K>(const MeshDev&, const OpConsts&, double, const double*, const double*, double*, long, \
                            hipStream_t);                                                                       \
  template void launch_d_minus_c<K>(const MeshDev&, const OpConsts&, double, const double*, const double*,         \
                                    const double*, const double*, double*, long, hipStream_t);                   \
  template void launch_d_minus_c_dots<K>(const MeshDev&, const OpConsts&, double, const double*, const double*,    \
                                         const double*, const double*, double*, long, const double*, long, int,  \
                                         double*, long, int, hipStream_t);                                       \
  template void launch_rhs_rho<K>(const MeshDev&, const OpConsts&, const Tableau&, int, const double* const*,     \
                                  double*, long, hipStream_t);                                                  \
  template void launch_rhs_pressure<K>(const MeshDev&, const OpConsts&, const Tableau&, int,                     \
                                       const double* const*, double*, long, hipStream_t);                       \
  template void launch_rhs_velocity<K>(const MeshDev&, const OpConsts&, const Tableau&, int,                     \
                                       const double* const*, double*, long, hipStream_t);

ATMO_INST_K(1)
ATMO_INST_K(2)
ATMO_INST_K(3)

}  // namespace atmo
