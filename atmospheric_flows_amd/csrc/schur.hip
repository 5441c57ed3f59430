// Cellwise-direct A^{-1} kernels (schur.h).
//
// The inverse blocks are stored per cell as ND2 x ND2 FP64 matrices. Dof a
// of a cell is produced by the thread (cell, a): it walks row a (= column a,
// the blocks are symmetric) while the 16 threads of the cell read 16
// consecutive doubles per step, so every load instruction moves whole
// 128-byte lines.
#include "cellface.h"
#include "schur.h"

namespace atmo {

template <int K, bool AMR>
__global__ __launch_bounds__(64) void k_b_ainv(MeshDev m, PhysB ph, const double* __restrict__ p,
                                               const double* __restrict__ Ainv, double* __restrict__ y, long Ns) {
  constexpr int ND2 = Cfg<K>::ND2, CPW = Cfg<K>::CPW, CS = lds_cell_stride<K>();
  __shared__ CFSmem<K, PhysB, AMR> sh;
  __shared__ double s_t[2][CPW * CS];
  FieldPtrs<1> in{{p}};
  cellface_columns<K, PhysB, AMR>(m, ph, in, sh);
  const int cell0 = blockIdx.x * CPW;
  OutRows<K> R = out_rows<K>(threadIdx.x);
  for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
    if (!rows_invariant<K>()) R = out_rows<K>(t);
    const int st = (t / ND2) * CS + t % ND2;
    s_t[0][st] = cellface_out<K, PhysB, AMR>(sh, 0, t, R);
    s_t[1][st] = cellface_out<K, PhysB, AMR>(sh, 1, t, R);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
    const int cw = t / ND2, a = t - cw * ND2;
    const long cell = cell0 + cw;
    if (cell >= m.ncell) break;
    const double* col = Ainv + cell * ND2 * ND2 + a;
    double y0 = 0.0, y1 = 0.0;
#pragma unroll
    for (int b = 0; b < ND2; ++b) {
      const double v = col[b * ND2];
      y0 += v * s_t[0][cw * CS + b];
      y1 += v * s_t[1][cw * CS + b];
    }
    const long gi = cell * ND2 + a;
    y[gi] = y0;
    y[Ns + gi] = y1;
  }
}

template <int K>
__global__ __launch_bounds__(256) void k_ainv_apply(int ncell, const double* __restrict__ Ainv,
                                                    const double* __restrict__ b, double* __restrict__ x, long Ns) {
  constexpr int ND2 = (K + 1) * (K + 1), CB = 256 / ND2;
  __shared__ double s_b[2][CB * ND2];
  const long base = (long)blockIdx.x * CB * ND2;
  const long lim = (long)ncell * ND2;
  const int t = threadIdx.x;
  if (t < CB * ND2) {
    const long gi = base + t;
    s_b[0][t] = gi < lim ? b[gi] : 0.0;
    s_b[1][t] = gi < lim ? b[Ns + gi] : 0.0;
  }
  __syncthreads();
  if (t >= CB * ND2) return;
  const int cw = t / ND2, a = t - cw * ND2;
  const long cell = (long)blockIdx.x * CB + cw;
  if (cell >= ncell) return;
  const double* col = Ainv + cell * ND2 * ND2 + a;
  double x0 = 0.0, x1 = 0.0;
#pragma unroll
  for (int k = 0; k < ND2; ++k) {
    const double v = col[k * ND2];
    x0 += v * s_b[0][cw * ND2 + k];
    x1 += v * s_b[1][cw * ND2 + k];
  }
  const long gi = cell * ND2 + a;
  x[gi] = x0;
  x[Ns + gi] = x1;
}

template <int K>
void launch_b_ainv(const MeshDev& m, const OpConsts& oc, double coeff, const double* p, const double* Ainv,
                   double* y, long Ns, hipStream_t s) {
  const int grid = ceil_div(m.ncell, Cfg<K>::CPW);
  if (grid == 0) return;
  PhysB ph;
  ph.cf = coeff * oc.dt / (oc.Ma * oc.Ma);
  if (m.cface)
    hipLaunchKernelGGL((k_b_ainv<K, true>), dim3(grid), dim3(64), 0, s, m, ph, p, Ainv, y, Ns);
  else
    hipLaunchKernelGGL((k_b_ainv<K, false>), dim3(grid), dim3(64), 0, s, m, ph, p, Ainv, y, Ns);
  ATMO_HIP_CHECK(hipGetLastError());
}

template <int K>
void launch_ainv_apply(const MeshDev& m, const double* Ainv, const double* b, double* x, long Ns, hipStream_t s) {
  constexpr int CB = 256 / ((K + 1) * (K + 1));
  const int grid = ceil_div(m.ncell, CB);
  if (grid == 0) return;
  hipLaunchKernelGGL((k_ainv_apply<K>), dim3(grid), dim3(256), 0, s, m.ncell, Ainv, b, x, Ns);
  ATMO_HIP_CHECK(hipGetLastError());
}

#define ATMO_SCHUR_INST(K)                                                                                     \
  template void launch_b_ainv<K>(const MeshDev&, const OpConsts&, double, const double*, const double*, double*, \
                                 long, hipStream_t);                                                          \
  template void launch_ainv_apply<K>(const MeshDev&, const double*, const double*, double*, long, hipStream_t);
ATMO_SCHUR_INST(1)
ATMO_SCHUR_INST(2)
ATMO_SCHUR_INST(3)

}  // namespace atmo
