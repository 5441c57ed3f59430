// Cellwise-direct A^{-1} kernels (schur.h).
//
// The inverse blocks are symmetric and stored packed (SymBlock: the upper
// triangle, N (N + 1) / 2 doubles per cell for N = (k+1)^2). A block's
// packed entries are loaded coalesced into LDS; dof a of a cell is then
// produced by the thread (cell, a) walking row a of the block in LDS.
#include "cellface.h"
#include "schur.h"

namespace atmo {

// The packed inverse blocks of nc consecutive cells, loaded coalesced into
// registers (load) and written to LDS (store: cell w at w * LDS_STRIDE; the
// cells [c0, c1) only, cell w at (w - c0) * LDS_STRIDE).
template <int K, int NT, int MAXC>
struct SymBlockStage {
  using SB = SymBlock<K>;
  static constexpr int NPASS = (MAXC * SB::SIZE + NT - 1) / NT;
  double v[NPASS];
  __device__ void load(const double* __restrict__ Ab, int nc) {
    const int lim = nc * SB::SIZE;
#pragma unroll
    for (int it = 0; it < NPASS; ++it) {
      const int e = threadIdx.x + it * NT;
      v[it] = e < lim ? Ab[e] : 0.0;
    }
  }
  __device__ void store(double* __restrict__ lds, int nc) const { store_cells(lds, 0, nc); }
  __device__ void store_cells(double* __restrict__ lds, int c0, int c1) const {
#pragma unroll
    for (int it = 0; it < NPASS; ++it) {
      const int e = threadIdx.x + it * NT;
      const int w = e / SB::SIZE;
      if (w >= c0 && w < c1) lds[e - c0 * SB::SIZE + (w - c0) * (SB::LDS_STRIDE - SB::SIZE)] = v[it];
    }
  }
};

// y = A^{-1} B p for the block's cells. The inverse blocks of the first half
// of the cells are loaded while B p is formed, those of the second half
// once it is (in flight during the first half's products); B p waits in
// registers until the cell+face partials are dead, and it and the inverse
// blocks of one half at a time take their LDS (one 9-cell block's cell+face
// LDS, not that plus the 9 inverse blocks, so more blocks fit a CU).
template <int K, bool AMR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(AMR ? 1 : 4))) void k_b_ainv(MeshDev m, PhysB ph, const double* __restrict__ p,
                                               const double* __restrict__ Ainv, double* __restrict__ y, long Ns) {
  constexpr int ND2 = Cfg<K>::ND2, CPW = Cfg<K>::CPW, CS = lds_cell_stride<K>();
  constexpr int NPASS = (CPW * ND2 + 63) / 64, HC = CPW / 2;  // cells of the first half
  constexpr int HMAX = CPW - HC > HC ? CPW - HC : HC;
  using SB = SymBlock<K>;
  __shared__ union {
    CFSmem<K, PhysB, AMR> cf;
    struct {
      double t[2][CPW * CS];        // B p
      double A[HMAX * SB::LDS_STRIDE];  // the inverse blocks of one half
    } ap;
  } sh;
  FieldPtrs<1> in{{p}};
  const int cell0 = cell_block() * CPW;
  const int nc = min(CPW, m.ncell - cell0);
  // the first half's inverse blocks are loaded with the column phase, the second half's after B p
  SymBlockStage<K, 64, HC> blk0;
  SymBlockStage<K, 64, CPW - HC> blk1;
  cellface_columns<K, PhysB, AMR>(m, ph, in, sh.cf,
                                  [&]() { blk0.load(Ainv + (long)cell0 * SB::SIZE, min(HC, nc)); });
  OutRows<K> R = out_rows<K>(threadIdx.x);
  double t0[NPASS], t1[NPASS];
#pragma unroll
  for (int ps = 0; ps < NPASS; ++ps) {
    const int t = threadIdx.x + 64 * ps;
    t0[ps] = t1[ps] = 0.0;
    if (t < CPW * ND2) {
      if (!rows_invariant<K>()) R = out_rows<K>(t);
      t0[ps] = cellface_out<K, PhysB, AMR>(sh.cf, 0, t, R);
      t1[ps] = cellface_out<K, PhysB, AMR>(sh.cf, 1, t, R);
    }
  }
  blk1.load(Ainv + (long)(cell0 + HC) * SB::SIZE, max(nc - HC, 0));
  __syncthreads();
#pragma unroll
  for (int ps = 0; ps < NPASS; ++ps) {
    const int t = threadIdx.x + 64 * ps;
    if (t < CPW * ND2) {
      const int st = (t / ND2) * CS + t % ND2;
      sh.ap.t[0][st] = t0[ps];
      sh.ap.t[1][st] = t1[ps];
    }
  }
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int c0 = half ? HC : 0, c1 = half ? nc : min(HC, nc);
    if (half) {
      __syncthreads();  // the first half's products have read their blocks
      blk1.store_cells(sh.ap.A, 0, c1 - c0);
    } else {
      blk0.store_cells(sh.ap.A, 0, c1);
    }
    __syncthreads();
#pragma unroll
    for (int ps = 0; ps < NPASS; ++ps) {
      const int t = threadIdx.x + 64 * ps;
      const int cw = t / ND2, a = t - cw * ND2;
      if (t >= CPW * ND2 || cw < c0 || cw >= c1) continue;
      const double* Ac = sh.ap.A + (cw - c0) * SB::LDS_STRIDE;
      double y0 = 0.0, y1 = 0.0;
#pragma unroll
      for (int b = 0; b < ND2; ++b) {
        const double v = Ac[SB::index(a, b)];
        y0 += v * sh.ap.t[0][cw * CS + b];
        y1 += v * sh.ap.t[1][cw * CS + b];
      }
      const long gi = (long)(cell0 + cw) * ND2 + a;
      y[gi] = y0;
      y[Ns + gi] = y1;
    }
  }
}

template <int K>
__global__ __launch_bounds__(256) void k_ainv_apply(int ncell, const double* __restrict__ Ainv,
                                                    const double* __restrict__ b, double* __restrict__ x, long Ns) {
  constexpr int ND2 = (K + 1) * (K + 1), CB = 256 / ND2, CS = lds_cell_stride<K>();
  using SB = SymBlock<K>;
  __shared__ double s_A[CB * SB::LDS_STRIDE];
  __shared__ double s_b[2][CB * CS];
  const long cell0 = (long)blockIdx.x * CB;
  const int nc = (int)min((long)CB, ncell - cell0);
  const int t = threadIdx.x;
  const int cw = t / ND2, a = t - cw * ND2;
  if (t < CB * ND2 && cw < nc) {
    const long gi = cell0 * ND2 + t;
    s_b[0][cw * CS + a] = b[gi];
    s_b[1][cw * CS + a] = b[Ns + gi];
  }
  SymBlockStage<K, 256, CB> blk;
  blk.load(Ainv + cell0 * SB::SIZE, nc);
  blk.store(s_A, nc);
  __syncthreads();
  if (t >= CB * ND2 || cw >= nc) return;
  const double* Ac = s_A + cw * SB::LDS_STRIDE;
  double x0 = 0.0, x1 = 0.0;
#pragma unroll
  for (int k = 0; k < ND2; ++k) {
    const double v = Ac[SB::index(a, k)];
    x0 += v * s_b[0][cw * CS + k];
    x1 += v * s_b[1][cw * CS + k];
  }
  const long gi = cell0 * ND2 + t;
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
