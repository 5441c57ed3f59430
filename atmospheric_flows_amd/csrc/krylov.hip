// Vector operations, order-fixed reductions and Krylov building blocks.
// CG follows deal.II SolverCG as used by the reference
// (Atmospheric_Flow.cc:452-536, euler_operator.h:2391-2397): residual
// g = A x - b, convergence on the unpreconditioned |g| <= tol, Jacobi
// preconditioner h = P g, direction d = -h + beta d.
#include "krylov.h"
#include "wave.h"

namespace atmo {

int red_blocks(long n) {
  long b = (n + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
  if (b < 1) b = 1;
  if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
  return (int)b;
}

// ------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// Block-level sum (or max) of NS values; result valid in thread 0.
template <int NS, bool MAX>
__device__ inline void block_reduce(double (&v)[NS]) {
  __shared__ double sh[NS][16];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    double x = MAX ? wave_max(v[k]) : wave_sum(v[k]);
    if (l == 0) sh[k][w] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double a = sh[k][0];
      for (int i = 1; i < nw; ++i) a = MAX ? fmax(a, sh[k][i]) : a + sh[k][i];
      v[k] = a;
    }
  }
  __syncthreads();
}

// Publish the block partials; the last block (ticket) reduces all partials
// in block order and calls fin(sum) in thread 0. Returns nothing; all
// threads of all blocks must call it.
template <int NS, bool MAX, class Fin>
__device__ inline void grid_reduce(double (&v)[NS], double* partials, unsigned* counter, Fin fin) {
  __shared__ int s_last;
  block_reduce<NS, MAX>(v);
  const int G = gridDim.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k)
      __hip_atomic_store(&partials[(long)k * RED_MAX_BLOCKS + blockIdx.x], v[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned t = atomicAdd(counter, 1u);
    s_last = (t == (unsigned)(G - 1));
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  double s[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) s[k] = MAX ? 0.0 : 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double p = __hip_atomic_load(&partials[(long)k * RED_MAX_BLOCKS + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s[k] = MAX ? fmax(s[k], p) : s[k] + p;
    }
  }
  block_reduce<NS, MAX>(s);
  if (threadIdx.x == 0) {
    fin(s);
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ------------------------------------------------------------ vector ops
template <typename T>
__global__ void k_axpby(long n, T a, const T* x, T b, T* y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (b == (T)0) ? a * x[i] : a * x[i] + b * y[i];
}
template <typename T>
__global__ void k_fill(long n, T v, T* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = v;
}
template <typename TO, typename TI>
__global__ void k_cast(long n, const TI* __restrict__ x, TO* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (TO)x[i];
}
template <typename T>
__global__ void k_mul(long n, const T* __restrict__ x, const T* __restrict__ d, T* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = x[i] * d[i];
}

static int ew_blocks(long n) {
  long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

template <typename T>
void vec_axpby(long n, T a, const T* x, T b, T* y, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby<T>, dim3(ew_blocks(n)), dim3(256), 0, s, n, a, x, b, y);
  ATMO_HIP_CHECK(hipGetLastError());
}
template <typename T>
void vec_fill(long n, T v, T* y, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<T>, dim3(ew_blocks(n)), dim3(256), 0, s, n, v, y);
  ATMO_HIP_CHECK(hipGetLastError());
}
template <typename T>
void vec_copy(long n, const T* x, T* y, hipStream_t s) {
  if (n <= 0 || x == y) return;
  ATMO_HIP_CHECK(hipMemcpyAsync(y, x, n * sizeof(T), hipMemcpyDeviceToDevice, s));
}
template <typename TO, typename TI>
void vec_cast(long n, const TI* x, TO* y, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_cast<TO, TI>), dim3(ew_blocks(n)), dim3(256), 0, s, n, x, y);
  ATMO_HIP_CHECK(hipGetLastError());
}
template <typename T>
void vec_mul(long n, const T* x, const T* d, T* y, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mul<T>, dim3(ew_blocks(n)), dim3(256), 0, s, n, x, d, y);
  ATMO_HIP_CHECK(hipGetLastError());
}

template void vec_axpby<double>(long, double, const double*, double, double*, hipStream_t);
template void vec_axpby<float>(long, float, const float*, float, float*, hipStream_t);
template void vec_fill<double>(long, double, double*, hipStream_t);
template void vec_fill<float>(long, float, float*, hipStream_t);
template void vec_copy<double>(long, const double*, double*, hipStream_t);
template void vec_copy<float>(long, const float*, float*, hipStream_t);
template void vec_cast<float, double>(long, const double*, float*, hipStream_t);
template void vec_cast<double, float>(long, const float*, double*, hipStream_t);
template void vec_mul<double>(long, const double*, const double*, double*, hipStream_t);
template void vec_mul<float>(long, const float*, const float*, float*, hipStream_t);

__global__ void k_dot(long n, const double* __restrict__ x, const double* __restrict__ y, double* out, int take_sqrt,
                      double* partials, unsigned* counter) {
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[0] += x[i] * y[i];
  grid_reduce<1, false>(v, partials, counter, [&](double (&s)[1]) { *out = take_sqrt ? sqrt(s[0]) : s[0]; });
}

__global__ void k_linf(long n, const double* __restrict__ x, const double* __restrict__ y, double* out,
                       double* partials, unsigned* counter) {
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[0] = fmax(v[0], fabs(y ? x[i] - y[i] : x[i]));
  grid_reduce<1, true>(v, partials, counter, [&](double (&s)[1]) { *out = s[0]; });
}

void vec_dot(long n, const double* x, const double* y, double* out, bool take_sqrt, RedBuf& rb, hipStream_t s) {
  hipLaunchKernelGGL(k_dot, dim3(red_blocks(n)), dim3(RED_THREADS), 0, s, n, x, y, out, take_sqrt ? 1 : 0,
                     rb.partials, rb.counter);
  ATMO_HIP_CHECK(hipGetLastError());
}

void vec_linf(long n, const double* x, const double* y, double* out, RedBuf& rb, hipStream_t s) {
  hipLaunchKernelGGL(k_linf, dim3(red_blocks(n)), dim3(RED_THREADS), 0, s, n, x, y, out, rb.partials, rb.counter);
  ATMO_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ generic CG
__device__ inline void cg_check(CGState* st, double res) {
  // SolverControl::check: success if res <= tol, failure if it >= max_its
  st->res = res;
  if (res <= st->tol) st->done = 1;
  else if (st->it >= st->max_its || isnan(res)) st->done = -1;
}

__global__ void k_cg_dot(long n, const double* __restrict__ x, const double* __restrict__ y, CGState* st, int mode,
                         double* partials, unsigned* counter) {
  if (st->done) return;
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[0] += x[i] * y[i];
  grid_reduce<1, false>(v, partials, counter, [&](double (&s)[1]) {
    if (mode == DOT_ALPHA) {
      st->dh = s[0];
      st->alpha = st->gh / s[0];
    } else if (mode == DOT_GH_INIT) {
      st->gh = s[0];
    } else if (mode == DOT_BETA) {
      st->beta = s[0] / st->gh;
      st->gh = s[0];
    }
  });
}

void cg_dot(long n, const double* x, const double* y, CGState* st, int mode, RedBuf& rb, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_dot, dim3(red_blocks(n)), dim3(RED_THREADS), 0, s, n, x, y, st, mode, rb.partials,
                     rb.counter);
  ATMO_HIP_CHECK(hipGetLastError());
}

__global__ void k_cg_update(long n, double* __restrict__ x, double* __restrict__ g, const double* __restrict__ d,
                            const double* __restrict__ h, CGState* st, double* partials, unsigned* counter) {
  if (st->done) return;
  const double a = st->alpha;
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    x[i] += a * d[i];
    const double gi = g[i] + a * h[i];
    g[i] = gi;
    v[0] += gi * gi;
  }
  grid_reduce<1, false>(v, partials, counter, [&](double (&s)[1]) {
    st->it += 1;
    cg_check(st, sqrt(s[0]));
  });
}

void cg_update(long n, double* x, double* g, const double* d, const double* h, CGState* st, RedBuf& rb,
               hipStream_t s) {
  hipLaunchKernelGGL(k_cg_update, dim3(red_blocks(n)), dim3(RED_THREADS), 0, s, n, x, g, d, h, st, rb.partials,
                     rb.counter);
  ATMO_HIP_CHECK(hipGetLastError());
}

__global__ void k_cg_dir(long n, double* __restrict__ d, const double* __restrict__ h, const CGState* st) {
  if (st->done) return;
  const double b = st->beta;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    d[i] = -h[i] + b * d[i];
}

void cg_dir(long n, double* d, const double* h, const CGState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_dir, dim3(ew_blocks(n)), dim3(256), 0, s, n, d, h, st);
  ATMO_HIP_CHECK(hipGetLastError());
}

__global__ void k_cg_residual(long n, double* __restrict__ g, const double* __restrict__ b, int have_ax) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    g[i] = have_ax ? g[i] - b[i] : -b[i];
}

void cg_residual(long n, double* g, const double* b, bool have_ax, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_residual, dim3(ew_blocks(n)), dim3(256), 0, s, n, g, b, have_ax ? 1 : 0);
  ATMO_HIP_CHECK(hipGetLastError());
}

__global__ void k_cg_init_check(long n, const double* __restrict__ g, CGState* st, int tol_from_res0,
                                double* partials, unsigned* counter) {
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[0] += g[i] * g[i];
  grid_reduce<1, false>(v, partials, counter, [&](double (&s)[1]) {
    const double r = sqrt(s[0]);
    st->res0 = r;
    if (tol_from_res0) st->tol = st->tol_factor * r;
    st->it = 0;
    st->done = 0;
    st->beta = 0;
    cg_check(st, r);
  });
}

void cg_init_check(long n, const double* g, CGState* st, bool tol_from_res0, RedBuf& rb, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_init_check, dim3(red_blocks(n)), dim3(RED_THREADS), 0, s, n, g, st,
                     tol_from_res0 ? 1 : 0, rb.partials, rb.counter);
  ATMO_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------ fused Jacobi CG
__global__ void k_jcg_init(long n, const double* __restrict__ b, const double* __restrict__ P, double* __restrict__ g,
                           double* __restrict__ d, CGState* st, double* partials, unsigned* counter) {
  double v[2] = {0.0, 0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double gi = -b[i];
    g[i] = gi;
    d[i] = 0.0;
    v[0] += gi * gi;
    v[1] += gi * P[i] * gi;
  }
  grid_reduce<2, false>(v, partials, counter, [&](double (&s)[2]) {
    const double r = sqrt(s[0]);
    st->res0 = r;
    st->tol = st->tol_factor * r;
    st->it = 0;
    st->done = 0;
    st->beta = 0;
    st->gh = s[1];
    cg_check(st, r);
  });
}

void jcg_init(long n, const double* b, const double* P, double* g, double* d, CGState* st, RedBuf& rb,
              hipStream_t s) {
  hipLaunchKernelGGL(k_jcg_init, dim3(red_blocks(n)), dim3(RED_THREADS), 0, s, n, b, P, g, d, st, rb.partials,
                     rb.counter);
  ATMO_HIP_CHECK(hipGetLastError());
}

// d = -P g + beta d (cell-local) ; h = A d ; partial d.h ; alpha = gh / dh
template <int K>
__global__ __launch_bounds__(64) void k_jcg_apply(MeshDev m, const double* __restrict__ rho,
                                                  const double* __restrict__ P, const double* __restrict__ g,
                                                  double* __restrict__ d, double* __restrict__ h, CGState* st,
                                                  long Ns, double* partials, unsigned* counter) {
  using C = Cfg<K>;
  constexpr int ND2 = C::ND2, CPW = C::CPW;
  if (st->done) return;
  __shared__ double s_u[2][CPW * ND2];
  __shared__ double s_r[CPW * ND2];
  __shared__ double s_red[2][CPW][C::ND][C::NQ];
  const double beta = st->beta;
  const long base = (long)blockIdx.x * CPW * ND2;
  const long lim = (long)m.ncell * ND2;
  for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
    const long gi = base + t;
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      double dv = 0;
      if (gi < lim) {
        const long k = comp * Ns + gi;
        dv = -P[k] * g[k] + beta * d[k];
        d[k] = dv;
      }
      s_u[comp][t] = dv;
    }
    s_r[t] = gi < lim ? rho[gi] : 0.0;
  }
  __syncthreads();
  wave_mass_core<K, double, 2, true>(m, blockIdx.x * CPW, &s_u[0][0], s_r, 1.0, &s_red[0][0][0][0]);
  __syncthreads();
  double v[1] = {0.0};
  for (int t = threadIdx.x; t < CPW * ND2; t += 64) {
    const long gi = base + t;
    if (gi >= lim) break;
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      const double hv = wave_mass_out<K, double>(&s_red[comp][0][0][0], t);
      h[comp * Ns + gi] = hv;
      v[0] += s_u[comp][t] * hv;
    }
  }
  grid_reduce<1, false>(v, partials, counter, [&](double (&s)[1]) {
    st->dh = s[0];
    st->alpha = st->gh / s[0];
  });
}

template <int K>
void jcg_apply(const MeshDev& m, const double* rho, const double* P, const double* g, double* d, double* h,
               CGState* st, RedBuf& rb, long Ns, hipStream_t s) {
  const int grid = ceil_div(m.ncell, Cfg<K>::CPW);
  if (grid > RED_MAX_BLOCKS * 64) throw std::invalid_argument("jcg_apply: grid too large");
  // The fused kernel reduces over its own grid; the partial buffer holds up
  // to RED_MAX_BLOCKS*64 blocks for slot 0 only.
  hipLaunchKernelGGL((k_jcg_apply<K>), dim3(grid), dim3(64), 0, s, m, rho, P, g, d, h, st, Ns, rb.partials,
                     rb.counter);
  ATMO_HIP_CHECK(hipGetLastError());
}

__global__ void k_jcg_update(long n, double* __restrict__ x, double* __restrict__ g, const double* __restrict__ d,
                             const double* __restrict__ h, const double* __restrict__ P, CGState* st,
                             double* partials, unsigned* counter) {
  if (st->done) return;
  const double a = st->alpha;
  double v[2] = {0.0, 0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    x[i] += a * d[i];
    const double gi = g[i] + a * h[i];
    g[i] = gi;
    v[0] += gi * gi;
    v[1] += gi * P[i] * gi;
  }
  grid_reduce<2, false>(v, partials, counter, [&](double (&s)[2]) {
    st->it += 1;
    cg_check(st, sqrt(s[0]));
    st->beta = s[1] / st->gh;
    st->gh = s[1];
  });
}

void jcg_update(long n, double* x, double* g, const double* d, const double* h, const double* P, CGState* st,
                RedBuf& rb, hipStream_t s) {
  hipLaunchKernelGGL(k_jcg_update, dim3(red_blocks(n)), dim3(RED_THREADS), 0, s, n, x, g, d, h, P, st, rb.partials,
                     rb.counter);
  ATMO_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ GMRES
__global__ void k_mdot(long n, const double* __restrict__ w, const double* __restrict__ V, long ldv, int nv,
                       double* out, double* partials, unsigned* counter) {
  // one pass over w per vector group of 8 to bound registers
  __shared__ int s_last;
  double acc[8];
  for (int g0 = 0; g0 < nv; g0 += 8) {
    const int ng = (nv - g0) < 8 ? (nv - g0) : 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
      const double wi = w[i];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < ng) acc[k] += wi * V[(long)(g0 + k) * ldv + i];
    }
    block_reduce<8, false>(acc);
    if (threadIdx.x == 0)
      for (int k = 0; k < ng; ++k)
        __hip_atomic_store(&partials[(long)(g0 + k) * RED_MAX_BLOCKS + blockIdx.x], acc[k], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned t = atomicAdd(counter, 1u);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  for (int g0 = 0; g0 < nv; g0 += 8) {
    const int ng = (nv - g0) < 8 ? (nv - g0) : 8;
    double s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < ng)
          s[k] += __hip_atomic_load(&partials[(long)(g0 + k) * RED_MAX_BLOCKS + b], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    block_reduce<8, false>(s);
    if (threadIdx.x == 0)
      for (int k = 0; k < ng; ++k) out[g0 + k] = s[k];
  }
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

void gm_mdot(long n, const double* w, const double* V, long ldv, int nv, double* out, RedBuf& rb, hipStream_t s) {
  if (nv > RED_MAX_SLOTS) throw std::invalid_argument("gm_mdot: too many vectors");
  hipLaunchKernelGGL(k_mdot, dim3(red_blocks(n)), dim3(RED_THREADS), 0, s, n, w, V, ldv, nv, out, rb.partials,
                     rb.counter);
  ATMO_HIP_CHECK(hipGetLastError());
}

__global__ void k_maxpy(long n, double* __restrict__ w, const double* __restrict__ V, long ldv, int nv,
                        const double* __restrict__ coef, double sign) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double a = w[i];
    for (int k = 0; k < nv; ++k) a += sign * coef[k] * V[(long)k * ldv + i];
    w[i] = a;
  }
}

void gm_maxpy(long n, double* w, const double* V, long ldv, int nv, const double* coef, hipStream_t s) {
  hipLaunchKernelGGL(k_maxpy, dim3(ew_blocks(n)), dim3(256), 0, s, n, w, V, ldv, nv, coef, -1.0);
  ATMO_HIP_CHECK(hipGetLastError());
}

void gm_update(long n, double* x, const double* V, long ldv, int nv, const double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_maxpy, dim3(ew_blocks(n)), dim3(256), 0, s, n, x, V, ldv, nv, y, 1.0);
  ATMO_HIP_CHECK(hipGetLastError());
}

__global__ void k_scale_inv(long n, const double* __restrict__ x, const double* __restrict__ sc,
                            double* __restrict__ y) {
  const double f = 1.0 / sc[0];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = x[i] * f;
}

void vec_scale_inv(long n, const double* x, const double* scal, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_inv, dim3(ew_blocks(n)), dim3(256), 0, s, n, x, scal, y);
  ATMO_HIP_CHECK(hipGetLastError());
}

template void jcg_apply<1>(const MeshDev&, const double*, const double*, const double*, double*, double*, CGState*,
                           RedBuf&, long, hipStream_t);
template void jcg_apply<2>(const MeshDev&, const double*, const double*, const double*, double*, double*, CGState*,
                           RedBuf&, long, hipStream_t);
template void jcg_apply<3>(const MeshDev&, const double*, const double*, const double*, double*, double*, CGState*,
                           RedBuf&, long, hipStream_t);

}  // namespace atmo
