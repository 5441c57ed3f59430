// Per-step diagnostics (Atmospheric_Flow.cc:991-1087): one thread per
// (cell, dof). The reference evaluates the fields at the QGaussLobatto(k+1)
// points (Atmospheric_Flow.cc:86-88, 233-235), which are FE_DGQ's support
// points, so each point value is the nodal value of the same index and the
// nodal L-infinity norm comes from the same pass. Maxima and minima skip
// non-finite values and count them instead, so a NaN in the state shows up
// as a nonzero count rather than vanishing in fmax/fmin. Block partials are
// combined by one finalize block; the result stays on the device until the
// caller reads the record.
#include <cmath>
#include <limits>

#include "diag.h"

namespace atmo {

namespace {

constexpr int DT = 256;  // threads per block

__device__ inline double wmax(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}
__device__ inline double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block max of a, b and block sum of c; thread 0 gets the results.
__device__ inline void block_max2_sum(double& a, double& b, double& c) {
  __shared__ double sh[3][DT / 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  a = wmax(a);
  b = wmax(b);
  c = wsum(c);
  if (l == 0) {
    sh[0][w] = a;
    sh[1][w] = b;
    sh[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < DT / 64; ++i) {
      a = fmax(a, sh[0][i]);
      b = fmax(b, sh[1][i]);
      c += sh[2][i];
    }
  }
}

// partials: [0,nb) = max of -rho at the points (the min, negated),
// [nb,2nb) = max |rho| at the dofs, [2nb,3nb) = non-finite count
__global__ void __launch_bounds__(DT) k_diag_density(const double* __restrict__ rho, long items,
                                                     double* __restrict__ part) {
  double nmin = -std::numeric_limits<double>::max(), amax = 0.0, bad = 0.0;
  for (long t = (long)blockIdx.x * DT + threadIdx.x; t < items; t += (long)gridDim.x * DT) {
    const double d = rho[t];
    // the point value (min) and the dof value (linfty_norm) are the same number: counted once per use
    if (isfinite(d)) {
      nmin = fmax(nmin, -d);
      amax = fmax(amax, fabs(d));
    } else {
      bad += 2.0;
    }
  }
  block_max2_sum(nmin, amax, bad);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = nmin;
    part[gridDim.x + blockIdx.x] = amax;
    part[2 * gridDim.x + blockIdx.x] = bad;
  }
}

// partials: [0,nb) = max celerity, [nb,2nb) = max |u|, [2nb,3nb) = non-finite count
__global__ void __launch_bounds__(DT) k_diag_flow(const double* __restrict__ rho, const double* __restrict__ u,
                                                  const double* __restrict__ p, double gamma_gas, long Ns,
                                                  double* __restrict__ part) {
  double cmax = 0.0, vmax = 0.0, bad = 0.0;
  for (long t = (long)blockIdx.x * DT + threadIdx.x; t < Ns; t += (long)gridDim.x * DT) {
    const double r = rho[t], pp = p[t], u0 = u[t], u1 = u[Ns + t];
    const double c = sqrt(gamma_gas * pp / r), w = sqrt(u0 * u0 + u1 * u1);
    if (isfinite(c)) cmax = fmax(cmax, c); else bad += 1.0;
    if (isfinite(w)) vmax = fmax(vmax, w); else bad += 1.0;
  }
  block_max2_sum(cmax, vmax, bad);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = cmax;
    part[gridDim.x + blockIdx.x] = vmax;
    part[2 * gridDim.x + blockIdx.x] = bad;
  }
}

// rec[s0] = sign * max(part[0,nb)), rec[s1] = max(part[nb,2nb)), rec[s2] = sum(part[2nb,3nb))
__global__ void __launch_bounds__(DT) k_diag_fin(const double* __restrict__ part, int nb, double sign0,
                                                 double* __restrict__ rec, int s0, int s1, int s2) {
  double a = -std::numeric_limits<double>::max(), b = -std::numeric_limits<double>::max(), c = 0.0;
  for (int i = threadIdx.x; i < nb; i += DT) {
    a = fmax(a, part[i]);
    b = fmax(b, part[nb + i]);
    c += part[2 * nb + i];
  }
  block_max2_sum(a, b, c);
  if (threadIdx.x == 0) {
    rec[s0] = sign0 * a;
    rec[s1] = b;
    rec[s2] = c;
  }
}

int diag_blocks(long items) {
  long b = (items + DT - 1) / DT;
  if (b > DIAG_MAX_BLOCKS) b = DIAG_MAX_BLOCKS;
  return b < 1 ? 1 : (int)b;
}

}  // namespace

void launch_diag_density_n(long items, const double* rho, int stage, double* partials, double* rec, hipStream_t s) {
  const int nb = diag_blocks(items);
  hipLaunchKernelGGL(k_diag_density, dim3(nb), dim3(DT), 0, s, rho, items, partials);
  hipLaunchKernelGGL(k_diag_fin, dim3(1), dim3(DT), 0, s, partials, nb, -1.0, rec, DG_MIN_RHO + stage,
                     DG_MAX_RHO + stage, DG_NONFINITE + stage);
}

template <int K>
void launch_diag_density(const MeshDev& m, const double* rho, int stage, double* partials, double* rec,
                         hipStream_t s) {
  const long items = (long)m.ncell * (K + 1) * (K + 1);
  const int nb = diag_blocks(items);
  hipLaunchKernelGGL(k_diag_density, dim3(nb), dim3(DT), 0, s, rho, items, partials);
  hipLaunchKernelGGL(k_diag_fin, dim3(1), dim3(DT), 0, s, partials, nb, -1.0, rec, DG_MIN_RHO + stage,
                     DG_MAX_RHO + stage, DG_NONFINITE + stage);
}

template <int K>
void launch_diag_flow(const MeshDev& m, const double* rho, const double* u, const double* p, double gamma_gas,
                      long Ns, double* partials, double* rec, hipStream_t s) {
  const int nb = diag_blocks(Ns);
  hipLaunchKernelGGL(k_diag_flow, dim3(nb), dim3(DT), 0, s, rho, u, p, gamma_gas, Ns, partials);
  hipLaunchKernelGGL(k_diag_fin, dim3(1), dim3(DT), 0, s, partials, nb, 1.0, rec, DG_MAX_CEL, DG_MAX_VEL,
                     DG_NONFINITE + 3);
}

double min_cell_diameter(int nx, int nz, double Lx, double H, const double* h) {
  double d = std::numeric_limits<double>::max();
  const double dx = Lx / nx;
  for (int j = 0; j < nz; ++j)
    for (int i = 0; i < nx; ++i) {
      const double z00 = h[i] + (H - h[i]) * j / nz, z10 = h[i + 1] + (H - h[i + 1]) * j / nz;
      const double z01 = h[i] + (H - h[i]) * (j + 1) / nz, z11 = h[i + 1] + (H - h[i + 1]) * (j + 1) / nz;
      const double d1 = std::hypot(dx, z11 - z00), d2 = std::hypot(dx, z01 - z10);
      d = std::min(d, std::max(d1, d2));
    }
  return d;
}

#define ATMO_DIAG_INST(K)                                                                                   \
  template void launch_diag_density<K>(const MeshDev&, const double*, int, double*, double*, hipStream_t); \
  template void launch_diag_flow<K>(const MeshDev&, const double*, const double*, const double*, double,   \
                                    long, double*, double*, hipStream_t);
ATMO_DIAG_INST(1)
ATMO_DIAG_INST(2)
ATMO_DIAG_INST(3)

}  // namespace atmo
