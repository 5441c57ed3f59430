// Vector operations, order-fixed reductions and Krylov building blocks.
// CG follows deal.II SolverCG as used by the reference
// (Atmospheric_Flow.cc:452-536, euler_operator.h:2391-2397): residual
// g = A x - b, convergence on the unpreconditioned |g| <= tol, preconditioned
// residual z = P g, direction d = -z + beta d, beta = (g.z)_new / (g.z)_old.
#include "krylov.h"
#include "wave.h"

namespace atmo {

int red_blocks(long n) {
  long b = (n + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
  if (b < 1) b = 1;
  if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
  return (int)b;
}

static int ew_blocks(long n) {
  long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
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

// Block-level sum (or max) of NS values, result broadcast to all threads.
// Must be called by every thread of the block.
template <int NS, bool MAX>
__device__ inline void block_reduce(double (&v)[NS]) {
  __shared__ double sh[NS][16];
  __shared__ double res[NS];
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
      res[k] = a;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NS; ++k) v[k] = res[k];
  __syncthreads();
}

template <int NS, bool MAX>
__device__ inline void write_partials(double (&v)[NS], double* part) {
  block_reduce<NS, MAX>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NS; ++k) part[(long)k * RED_MAX_BLOCKS + blockIdx.x] = v[k];
}

// sum of G partials (one slot), identical in every block (fixed order);
// blockDim.x = 256, every load issued before the first add
__device__ inline double sum_slot(const double* part, int G) {
  constexpr int NB = RED_MAX_BLOCKS / RED_THREADS;
  double p[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = threadIdx.x + RED_THREADS * i;
    p[i] = b < G ? part[b] : 0.0;
  }
  double v[1] = {p[0]};
#pragma unroll
  for (int i = 1; i < NB; ++i) v[0] += p[i];
  block_reduce<1, false>(v);
  return v[0];
}

__device__ inline void cg_check(CGState* st, double res) {
  // SolverControl::check: success if res <= tol, failure if it >= max_its
  st->res = res;
  if (res <= st->tol) st->done = 1;
  else if (st->it >= st->max_its || isnan(res)) st->done = -1;
}

enum FinMode {
  FIN_DOT = 0, FIN_SQRT, FIN_MAX, FIN_CG_ALPHA, FIN_CG_GH, FIN_CG_BETA, FIN_CG_UPDATE, FIN_CG_INIT,
  FIN_CG_INIT_TOL, FIN_JCG_INIT, FIN_MDOT, FIN_GM_NORM
};

// One block: sums each of the nslot partial slots over its G blocks in
// block order (one wave per slot, lanes strided over the blocks, then a
// fixed shuffle tree) and finalizes. FIN_GM_NORM: *out = |v|, *out2 = 1/|v|
// (0 for a zero vector).
__global__ void k_fin(const double* __restrict__ part, int G, int nslot, int mode, double* out, CGState* st,
                      double* out2) {
  __shared__ double s_sum[RED_MAX_SLOTS];
  constexpr int NB = RED_MAX_BLOCKS / 64;  // partials per lane (fixed count: all loads in flight at once)
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int k = w; k < nslot; k += nw) {
    double p[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = l + 64 * i;
      p[i] = b < G ? part[(long)k * RED_MAX_BLOCKS + b] : 0.0;
    }
    double v = p[0];
#pragma unroll
    for (int i = 1; i < NB; ++i) v = mode == FIN_MAX ? fmax(v, p[i]) : v + p[i];
    v = mode == FIN_MAX ? wave_max(v) : wave_sum(v);
    if (l == 0) s_sum[k] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double s0 = s_sum[0];
  switch (mode) {
    case FIN_DOT: *out = s0; break;
    case FIN_SQRT: *out = sqrt(s0); break;
    case FIN_GM_NORM: {
      const double r = sqrt(s0);
      *out = r;
      *out2 = r > 0.0 ? 1.0 / r : 0.0;
      break;
    }
    case FIN_MAX: *out = s0; break;
    case FIN_CG_ALPHA:
      if (!st->done) {
        st->dh = s0;
        st->alpha = st->gh / s0;
        if (out) *out = st->alpha;
      }
      break;
    case FIN_CG_GH:
      if (!st->done) st->gh = s0;
      break;
    case FIN_CG_BETA:
      if (!st->done) {
        st->beta = s0 / st->gh;
        st->gh = s0;
        if (out) *out = st->beta;
      }
      break;
    case FIN_CG_UPDATE:
      if (!st->done) { st->it += 1; cg_check(st, sqrt(s0)); }
      break;
    case FIN_CG_INIT:
    case FIN_CG_INIT_TOL: {
      const double r = sqrt(s0);
      st->res0 = r;
      if (mode == FIN_CG_INIT_TOL) st->tol = st->tol_factor * r;
      st->it = 0;
      st->done = 0;
      st->beta = 0;
      cg_check(st, r);
      break;
    }
    case FIN_JCG_INIT: {
      const double r = sqrt(s0);
      st->res0 = r;
      st->tol = st->tol_factor * r;
      st->it = 0;
      st->done = 0;
      st->beta = 0;
      st->gh = s_sum[1];
      cg_check(st, r);
      break;
    }
    case FIN_MDOT:
      for (int k = 0; k < nslot; ++k) out[k] = s_sum[k];
      break;
  }
}

static void fin(const double* part, int G, int nslot, int mode, double* out, CGState* st, hipStream_t s,
                double* out2 = nullptr) {
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(RED_THREADS), 0, s, part, G, nslot, mode, out, st, out2);
  ATMO_HIP_CHECK(hipGetLastError());
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

__global__ void k_dot_part(long n, const double* __restrict__ x, const double* __restrict__ y, double* part,
                           const CGState* st) {
  if (st && st->done) return;
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[0] += x[i] * y[i];
  write_partials<1, false>(v, part);
}

__global__ void k_linf_part(long n, const double* __restrict__ x, const double* __restrict__ y, double* part) {
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[0] = fmax(v[0], fabs(y ? x[i] - y[i] : x[i]));
  write_partials<1, true>(v, part);
}

void vec_dot(long n, const double* x, const double* y, double* out, bool take_sqrt, RedBuf& rb, hipStream_t s) {
  const int G = red_blocks(n);
  hipLaunchKernelGGL(k_dot_part, dim3(G), dim3(RED_THREADS), 0, s, n, x, y, rb.gen, (const CGState*)nullptr);
  ATMO_HIP_CHECK(hipGetLastError());
  fin(rb.gen, G, 1, take_sqrt ? FIN_SQRT : FIN_DOT, out, nullptr, s);
}

void vec_linf(long n, const double* x, const double* y, double* out, RedBuf& rb, hipStream_t s) {
  const int G = red_blocks(n);
  hipLaunchKernelGGL(k_linf_part, dim3(G), dim3(RED_THREADS), 0, s, n, x, y, rb.gen);
  ATMO_HIP_CHECK(hipGetLastError());
  fin(rb.gen, G, 1, FIN_MAX, out, nullptr, s);
}

// ------------------------------------------------------------ generic CG
void cg_dot(long n, const double* x, const double* y, CGState* st, int mode, RedBuf& rb, hipStream_t s,
            double* record) {
  const int G = red_blocks(n);
  hipLaunchKernelGGL(k_dot_part, dim3(G), dim3(RED_THREADS), 0, s, n, x, y, rb.gen, (const CGState*)st);
  ATMO_HIP_CHECK(hipGetLastError());
  const int fm = mode == DOT_ALPHA ? FIN_CG_ALPHA : (mode == DOT_GH_INIT ? FIN_CG_GH : FIN_CG_BETA);
  fin(rb.gen, G, 1, fm, record, st, s);
}

__global__ void k_cg_update(long n, double* __restrict__ x, double* __restrict__ g, const double* __restrict__ d,
                            const double* __restrict__ h, CGState* st, double* part) {
  if (st->done) return;
  const double a = st->alpha;
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    x[i] += a * d[i];
    const double gi = g[i] + a * h[i];
    g[i] = gi;
    v[0] += gi * gi;
  }
  write_partials<1, false>(v, part);
}

void cg_update(long n, double* x, double* g, const double* d, const double* h, CGState* st, RedBuf& rb,
               hipStream_t s) {
  const int G = red_blocks(n);
  hipLaunchKernelGGL(k_cg_update, dim3(G), dim3(RED_THREADS), 0, s, n, x, g, d, h, st, rb.gen);
  ATMO_HIP_CHECK(hipGetLastError());
  fin(rb.gen, G, 1, FIN_CG_UPDATE, nullptr, st, s);
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

void cg_init_check(long n, const double* g, CGState* st, bool tol_from_res0, RedBuf& rb, hipStream_t s) {
  const int G = red_blocks(n);
  hipLaunchKernelGGL(k_dot_part, dim3(G), dim3(RED_THREADS), 0, s, n, g, g, rb.gen, (const CGState*)nullptr);
  ATMO_HIP_CHECK(hipGetLastError());
  fin(rb.gen, G, 1, tol_from_res0 ? FIN_CG_INIT_TOL : FIN_CG_INIT, nullptr, st, s);
}

// --------------------------------------------------- fused inner CG (A)
// Update kernel: one thread per (cell, dof), CB cells per block batch.
// INIT: g = -b (x = 0), d = 0. Otherwise x += alpha d, g += alpha h with
// alpha = gh / dh re-reduced from the partials of jcg_apply (pa) and of the
// previous update (pb[(i-1)&1]). Jacobi-preconditioned residual z = P g.
// Partials (g.g, g.z) -> pb[i&1].
template <int K, bool INIT>
__global__ __launch_bounds__(256) void k_jcg_upd(int i, long Ns, int ncell, double* __restrict__ x,
                                                 double* __restrict__ g, double* __restrict__ d,
                                                 const double* __restrict__ h, const double* __restrict__ b,
                                                 const double* __restrict__ P, CGState* st, double* pa, int Ga,
                                                 double* pb, int Gb) {
  constexpr int ND2 = (K + 1) * (K + 1);
  constexpr int CB = 256 / ND2;
  if (!INIT && st->done) return;
  double alpha = 0.0;
  if (!INIT) {
    const double dh = sum_slot(pa, Ga);
    const double gh = sum_slot(pb + (long)((i - 1) & 1) * 2 * RED_MAX_BLOCKS + RED_MAX_BLOCKS, Gb);
    alpha = gh / dh;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->alpha = alpha;
      st->dh = dh;
      st->it = i;
    }
  }
  const int cb = threadIdx.x / ND2, li = threadIdx.x - cb * ND2;
  const int nbatch = (ncell + CB - 1) / CB;
  double v[2] = {0.0, 0.0};
  for (int bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    const int cell = bt * CB + cb;
    if (cb >= CB || cell >= ncell) continue;
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      const long k = comp * Ns + (long)cell * ND2 + li;
      double gv;
      if (INIT) {
        gv = -b[k];
        d[k] = 0.0;
      } else {
        x[k] += alpha * d[k];
        gv = g[k] + alpha * h[k];
      }
      g[k] = gv;
      v[0] += gv * gv;
      v[1] += gv * P[k] * gv;
    }
  }
  block_reduce<2, false>(v);
  if (threadIdx.x == 0) {
    double* dst = pb + (long)(i & 1) * 2 * RED_MAX_BLOCKS;
    dst[blockIdx.x] = v[0];
    dst[RED_MAX_BLOCKS + blockIdx.x] = v[1];
  }
}

static int upd_blocks(int ncell, int K) {
  const int CB = 256 / ((K + 1) * (K + 1));
  long b = (ncell + CB - 1) / CB;
  if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
  return (int)std::max(1L, b);
}

template <int K>
static void launch_upd(bool init, int i, long Ns, const MeshDev& m, double* x, double* g, double* d,
                       const double* h, const double* b, const double* P, CGState* st, RedBuf& rb, hipStream_t s) {
  const int G = upd_blocks(m.ncell, K);
  const int Ga = std::min(RED_MAX_BLOCKS, std::max(1, ceil_div(ceil_div(m.ncell, Cfg<K>::CPW), 4)));
  if (init)
    hipLaunchKernelGGL((k_jcg_upd<K, true>), dim3(G), dim3(256), 0, s, i, Ns, m.ncell, x, g, d, h, b, P, st, rb.pa,
                       Ga, rb.pb, G);
  else
    hipLaunchKernelGGL((k_jcg_upd<K, false>), dim3(G), dim3(256), 0, s, i, Ns, m.ncell, x, g, d, h, b, P, st, rb.pa,
                       Ga, rb.pb, G);
  ATMO_HIP_CHECK(hipGetLastError());
}

void jcg_init(long Ns, const MeshDev& m, int K, const double* b, const double* P, double* g, double* d, CGState* st,
              RedBuf& rb, hipStream_t s) {
  if (K == 1) launch_upd<1>(true, 0, Ns, m, nullptr, g, d, nullptr, b, P, st, rb, s);
  else if (K == 2) launch_upd<2>(true, 0, Ns, m, nullptr, g, d, nullptr, b, P, st, rb, s);
  else launch_upd<3>(true, 0, Ns, m, nullptr, g, d, nullptr, b, P, st, rb, s);
  const int G = upd_blocks(m.ncell, K);
  fin(rb.pb, G, 2, FIN_JCG_INIT, nullptr, st, s);
}

void jcg_update(int i, long Ns, const MeshDev& m, int K, double* x, double* g, const double* d, const double* h,
                const double* P, CGState* st, RedBuf& rb, hipStream_t s) {
  double* dd = const_cast<double*>(d);
  if (K == 1) launch_upd<1>(false, i, Ns, m, x, g, dd, h, nullptr, P, st, rb, s);
  else if (K == 2) launch_upd<2>(false, i, Ns, m, x, g, dd, h, nullptr, P, st, rb, s);
  else launch_upd<3>(false, i, Ns, m, x, g, dd, h, nullptr, P, st, rb, s);
}

// Apply kernel of iteration i: (re-reducing the previous update's partials)
// convergence test on |g_{i-1}|, beta = gz_{i-1}/gz_{i-2}; then per cell
// d = -P g + beta d, h = A d, partial d.h -> pa.
// Four independent waves per block, each owning CPW cells at a time.
template <int K>
__global__ __launch_bounds__(256) void k_jcg_apply(int i, MeshDev m, const double* __restrict__ rho,
                                                   const double* __restrict__ P,
                                                   const double* __restrict__ g, double* __restrict__ d,
                                                   double* __restrict__ h, CGState* st, long Ns, double* pa,
                                                   const double* pb, int Gb) {
  using C = Cfg<K>;
  constexpr int ND2 = C::ND2, CPW = C::CPW, NQ = C::NQ, ND = C::ND;
  __shared__ double s_u[4][2][CPW * ND2];
  __shared__ double s_r[4][CPW * ND2];
  __shared__ double s_red[4][2][CPW * ND * NQ];
  __shared__ int s_done;
  if (st->done) return;
  double beta = 0.0;
  if (i >= 2) {
    const double* newer = pb + (long)((i - 1) & 1) * 2 * RED_MAX_BLOCKS;
    const double* older = pb + (long)(i & 1) * 2 * RED_MAX_BLOCKS;
    const double gg = sum_slot(newer, Gb);
    const double gz = sum_slot(newer + RED_MAX_BLOCKS, Gb);
    const double gz_old = sum_slot(older + RED_MAX_BLOCKS, Gb);
    const double res = sqrt(gg);
    if (threadIdx.x == 0) {
      int dn = 0;
      if (res <= st->tol) dn = 1;
      else if (i - 1 >= st->max_its || isnan(res)) dn = -1;
      s_done = dn;
      if (blockIdx.x == 0) {
        st->res = res;
        st->beta = gz / gz_old;
        st->gh = gz;
        if (dn) st->done = dn;
      }
    }
    __syncthreads();
    if (s_done) return;
    beta = gz / gz_old;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ngroups = (m.ncell + CPW - 1) / CPW;
  const long lim = (long)m.ncell * ND2;
  double acc = 0.0;
  for (int grp = blockIdx.x * 4 + w; grp < ngroups; grp += gridDim.x * 4) {
    const long base = (long)grp * CPW * ND2;
    for (int t = lane; t < CPW * ND2; t += 64) {
      const long gi = base + t;
#pragma unroll
      for (int comp = 0; comp < 2; ++comp) {
        double dv = 0.0;
        if (gi < lim) {
          const long k = comp * Ns + gi;
          dv = -P[k] * g[k] + beta * d[k];
          d[k] = dv;
        }
        s_u[w][comp][t] = dv;
      }
      s_r[w][t] = gi < lim ? rho[gi] : 0.0;
    }
    wave_sync();
    wave_mass_core<K, double, 2, true>(m, grp * CPW, &s_u[w][0][0], s_r[w], 1.0, &s_red[w][0][0]);
    wave_sync();
    double Si[NQ];
    lane_cols_S<K, double>(lane % ND, Si);
    for (int t = lane; t < CPW * ND2; t += 64) {
      const long gi = base + t;
      if (gi >= lim) break;
      if (64 % ND != 0) lane_cols_S<K, double>(t % ND, Si);
#pragma unroll
      for (int comp = 0; comp < 2; ++comp) {
        const double hv = wave_mass_out<K, double>(&s_red[w][comp][0], t, Si);
        h[comp * Ns + gi] = hv;
        acc += s_u[w][comp][t] * hv;
      }
    }
    wave_sync();
  }
  double v[1] = {acc};
  block_reduce<1, false>(v);
  if (threadIdx.x == 0) pa[blockIdx.x] = v[0];
}

template <int K>
void jcg_apply(int i, const MeshDev& m, const double* rho, const double* P, const double* g, double* d, double* h,
               CGState* st, RedBuf& rb, long Ns, hipStream_t s) {
  const int ngroups = ceil_div(m.ncell, Cfg<K>::CPW);
  const int Ga = std::min(RED_MAX_BLOCKS, std::max(1, ceil_div(ngroups, 4)));
  const int Gb = upd_blocks(m.ncell, K);
  hipLaunchKernelGGL((k_jcg_apply<K>), dim3(Ga), dim3(256), 0, s, i, m, rho, P, g, d, h, st, Ns, rb.pa, rb.pb, Gb);
  ATMO_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------ cell block inverses
// One wave per cell: assemble the (k+1)^2 x (k+1)^2 block of A (rho-weighted
// mass at the 2k+1 Gauss points) in LDS and invert it in place by
// Gauss-Jordan (SPD, no pivoting) in FP64. Stored packed (SymBlock), each
// off-diagonal pair as the mean of its two computed entries.
template <int K>
__global__ __launch_bounds__(256) void k_block_inv(MeshDev m, const double* __restrict__ rho,
                                                   double* __restrict__ Ainv) {
  using C = Cfg<K>;
  constexpr int ND = C::ND, ND2 = C::ND2, NQ = C::NQ, NQ2 = NQ * NQ;
  constexpr auto& tb = basis::kT<K>;
  __shared__ double s_A[4][ND2 * ND2];
  __shared__ double s_w[4][NQ2];
  __shared__ double s_S[NQ * ND];
  __shared__ double s_row[4][ND2], s_col[4][ND2];
  __shared__ double s_T[4][NQ * ND * ND];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int t = threadIdx.x; t < NQ * ND; t += blockDim.x) {
    double v = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int i = 0; i < ND; ++i)
        if (q * ND + i == t) v = tb.S[q][i];
    s_S[t] = v;
  }
  __syncthreads();
  const int cell = blockIdx.x * 4 + w;
  if (cell >= m.ncell) return;
  const CellVerts vv = cell_verts(m, cell);
  // weights JxW * rho at the NQ^2 points (lane = q)
  for (int q = lane; q < NQ2; q += 64) {
    const int c = q % NQ, r = q / NQ;
    const ColGeo<double> g = col_geo<K, double>(vv, c);
    double eta = 0, wr = 0;
#pragma unroll
    for (int rr = 0; rr < NQ; ++rr)
      if (rr == r) {
        eta = tb.gq[rr];
        wr = tb.gw[rr];
      }
    (void)eta;
    const double jxw = g.dx * g.ze * g.wc * wr;
    double rv = 0;
    for (int j = 0; j < ND; ++j)
      for (int i = 0; i < ND; ++i) rv += s_S[c * ND + i] * s_S[r * ND + j] * rho[(long)cell * ND2 + j * ND + i];
    s_w[w][q] = jxw * rv;
  }
  wave_sync();
  // sum factorization: T[r][ia][ib] = sum_c S[c][ia] S[c][ib] w[r][c], then
  // A[(ja,ia)][(jb,ib)] = sum_r S[r][ja] S[r][jb] T[r][ia][ib] (exactly symmetric)
  for (int e = lane; e < NQ * ND * ND; e += 64) {
    const int r = e / (ND * ND), pq = e - r * (ND * ND), ia = pq / ND, ib = pq - ia * ND;
    double t = 0;
#pragma unroll
    for (int c = 0; c < NQ; ++c) t += s_S[c * ND + ia] * s_S[c * ND + ib] * s_w[w][r * NQ + c];
    s_T[w][e] = t;
  }
  wave_sync();
  for (int e = lane; e < ND2 * ND2; e += 64) {
    const int a = e / ND2, b = e - a * ND2;
    const int ja = a / ND, ia = a - ja * ND, jb = b / ND, ib = b - jb * ND;
    double s = 0;
#pragma unroll
    for (int r = 0; r < NQ; ++r) s += s_S[r * ND + ja] * s_S[r * ND + jb] * s_T[w][(r * ND + ia) * ND + ib];
    s_A[w][e] = s;
  }
  wave_sync();
  for (int k = 0; k < ND2; ++k) {
    const double piv = s_A[w][k * ND2 + k];
    for (int j = lane; j < ND2; j += 64) {
      s_row[w][j] = (j == k ? 1.0 : s_A[w][k * ND2 + j]) / piv;
      s_col[w][j] = s_A[w][j * ND2 + k];
    }
    wave_sync();
    for (int e = lane; e < ND2 * ND2; e += 64) {
      const int r = e / ND2, j = e - r * ND2;
      if (r == k) s_A[w][e] = s_row[w][j];
      else s_A[w][e] = (j == k ? 0.0 : s_A[w][e]) - s_col[w][r] * s_row[w][j];
    }
    wave_sync();
  }
  for (int e = lane; e < ND2 * ND2; e += 64) {
    const int a = e / ND2, b = e - a * ND2;
    if (b >= a)
      Ainv[(long)cell * SymBlock<K>::SIZE + SymBlock<K>::index(a, b)] = 0.5 * (s_A[w][e] + s_A[w][b * ND2 + a]);
  }
}

template <int K>
void build_block_inverse(const MeshDev& m, const double* rho, double* Ainv, hipStream_t s) {
  hipLaunchKernelGGL((k_block_inv<K>), dim3(ceil_div(m.ncell, 4)), dim3(256), 0, s, m, rho, Ainv);
  ATMO_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ GMRES
// Partial dot products of w with nv columns of V, MDOT_GROUP columns per
// pass over w. PAIRS: two consecutive entries per thread and step (16-byte
// loads; needs n and ldv even, i.e. every column 16-byte aligned).
constexpr int MDOT_GROUP = 16;
template <bool PAIRS>
__global__ __launch_bounds__(RED_THREADS) void k_mdot_part(long n, const double* __restrict__ w,
                                                            const double* __restrict__ V, long ldv, int nv,
                                                            double* part) {
  double acc[MDOT_GROUP];
  const long nstep = PAIRS ? n / 2 : n;
  for (int g0 = 0; g0 < nv; g0 += MDOT_GROUP) {
    const int ng = (nv - g0) < MDOT_GROUP ? (nv - g0) : MDOT_GROUP;
#pragma unroll
    for (int k = 0; k < MDOT_GROUP; ++k) acc[k] = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nstep; i += (long)gridDim.x * blockDim.x) {
      if (PAIRS) {
        const double2 wi = reinterpret_cast<const double2*>(w)[i];
#pragma unroll
        for (int k = 0; k < MDOT_GROUP; ++k)
          if (k < ng) {
            const double2 v = reinterpret_cast<const double2*>(V + (long)(g0 + k) * ldv)[i];
            acc[k] += wi.x * v.x + wi.y * v.y;
          }
      } else {
        const double wi = w[i];
#pragma unroll
        for (int k = 0; k < MDOT_GROUP; ++k)
          if (k < ng) acc[k] += wi * V[(long)(g0 + k) * ldv + i];
      }
    }
    block_reduce<MDOT_GROUP, false>(acc);
    if (threadIdx.x == 0)
      for (int k = 0; k < ng; ++k) part[(long)(g0 + k) * RED_MAX_BLOCKS + blockIdx.x] = acc[k];
  }
}

void gm_mdot(long n, const double* w, const double* V, long ldv, int nv, double* out, RedBuf& rb, hipStream_t s) {
  if (nv > RED_MAX_SLOTS) throw std::invalid_argument("gm_mdot: too many vectors");
  const int G = red_blocks(n);
  const bool pairs = ((n | ldv) & 1) == 0 && ((reinterpret_cast<uintptr_t>(w) | reinterpret_cast<uintptr_t>(V)) & 15) == 0;
  if (pairs)
    hipLaunchKernelGGL(k_mdot_part<true>, dim3(G), dim3(RED_THREADS), 0, s, n, w, V, ldv, nv, rb.gen);
  else
    hipLaunchKernelGGL(k_mdot_part<false>, dim3(G), dim3(RED_THREADS), 0, s, n, w, V, ldv, nv, rb.gen);
  ATMO_HIP_CHECK(hipGetLastError());
  fin(rb.gen, G, nv, FIN_MDOT, out, nullptr, s);
}

__global__ void k_maxpy(long n, double* __restrict__ w, const double* __restrict__ V, long ldv, int nv,
                        const double* __restrict__ coef, double sign) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double a = w[i];
    for (int k = 0; k < nv; ++k) a += sign * coef[k] * V[(long)k * ldv + i];
    w[i] = a;
  }
}

// w -= sum_k c_k V_k with c_k = sig[k]^2 dots[k] (the projection on the
// unnormalized basis vectors V_k = v_k / sig[k]) and partial |w|^2 of the
// result in the same pass. PAIRS: two consecutive entries per thread and
// step (16-byte accesses); four independent partial sums over k keep
// several loads in flight.
// The basis columns are read with non-temporal loads (as in the Arnoldi kernel's dot products): streamed
// through once per use, they then do not displace the data the other kernels of the GMRES step re-read
// (measured: the mixed step 3 % faster with both kernels streaming, slower with either one alone).
typedef double nt_d2 __attribute__((ext_vector_type(2)));
template <bool PAIRS>
__global__ __launch_bounds__(RED_THREADS) void k_maxpy_norm(long n, double* __restrict__ w,
                                                             const double* __restrict__ V, long ldv, int nv,
                                                             const double* __restrict__ dots,
                                                             const double* __restrict__ sig, double* part) {
  constexpr int E = PAIRS ? 2 : 1;
  __shared__ double coef[RED_MAX_SLOTS];
  for (int k = threadIdx.x; k < nv; k += blockDim.x) coef[k] = sig[k] * sig[k] * dots[k];
  __syncthreads();
  double acc[1] = {0.0};
  const long nstep = n / E;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nstep; i += (long)gridDim.x * blockDim.x) {
    double a[4][E] = {};
    int k = 0;
    for (; k + 4 <= nv; k += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double* col = V + (long)(k + u) * ldv + i * E;
        const double c = coef[k + u];
        if (PAIRS) {
          const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(col));
          a[u][0] += c * v.x;
          a[u][E - 1] += c * v.y;
        } else {
          a[u][0] += c * __builtin_nontemporal_load(col);
        }
      }
    }
    for (; k < nv; ++k) {
      const double* col = V + (long)k * ldv + i * E;
#pragma unroll
      for (int e = 0; e < E; ++e) a[0][e] += coef[k] * col[e];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const double r = w[i * E + e] - ((a[0][e] + a[1][e]) + (a[2][e] + a[3][e]));
      w[i * E + e] = r;
      acc[0] += r * r;
    }
  }
  block_reduce<1, false>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

void gm_maxpy_norm(long n, double* w, const double* V, long ldv, int nv, const double* dots, const double* sig,
                   double* norm, double* inv_norm, RedBuf& rb, hipStream_t s) {
  if (nv > RED_MAX_SLOTS) throw std::invalid_argument("gm_maxpy_norm: too many vectors");
  const int G = red_blocks(n);
  const bool pairs = ((n | ldv) & 1) == 0 && ((reinterpret_cast<uintptr_t>(w) | reinterpret_cast<uintptr_t>(V)) & 15) == 0;
  if (pairs)
    hipLaunchKernelGGL(k_maxpy_norm<true>, dim3(G), dim3(RED_THREADS), 0, s, n, w, V, ldv, nv, dots, sig, rb.gen);
  else
    hipLaunchKernelGGL(k_maxpy_norm<false>, dim3(G), dim3(RED_THREADS), 0, s, n, w, V, ldv, nv, dots, sig, rb.gen);
  ATMO_HIP_CHECK(hipGetLastError());
  fin(rb.gen, G, 1, FIN_GM_NORM, norm, nullptr, s, inv_norm);
}

// out[s] = sum over b < G of part[s*pstride + b] for s < nslot: one block
// of 1024 threads per slot, fixed order (each thread's strided terms, then
// the block tree), every load of a round issued before the first add.
__global__ __launch_bounds__(1024) void k_sum_slots(const double* __restrict__ part, long pstride, int G,
                                                    double* __restrict__ out) {
  constexpr int U = 16;
  const double* p = part + (long)blockIdx.x * pstride;
  double acc[1] = {0.0};
  for (int b0 = threadIdx.x; b0 < G; b0 += 1024 * U) {
    double t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + 1024 * u;
      t[u] = b < G ? p[b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc[0] += t[u];
  }
  block_reduce<1, false>(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = acc[0];
}

void gm_sum_slots(const double* part, long pstride, int G, int nslot, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_slots, dim3(nslot), dim3(1024), 0, s, part, pstride, G, out);
  ATMO_HIP_CHECK(hipGetLastError());
}

void gm_norm(long n, const double* v, double* norm, double* inv_norm, RedBuf& rb, hipStream_t s) {
  const int G = red_blocks(n);
  hipLaunchKernelGGL(k_dot_part, dim3(G), dim3(RED_THREADS), 0, s, n, v, v, rb.gen, (const CGState*)nullptr);
  ATMO_HIP_CHECK(hipGetLastError());
  fin(rb.gen, G, 1, FIN_GM_NORM, norm, nullptr, s, inv_norm);
}

void gm_update(long n, double* x, const double* V, long ldv, int nv, const double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_maxpy, dim3(ew_blocks(n)), dim3(256), 0, s, n, x, V, ldv, nv, y, 1.0);
  ATMO_HIP_CHECK(hipGetLastError());
}

#define ATMO_KRY_INST(K)                                                                                          \
  template void jcg_apply<K>(int, const MeshDev&, const double*, const double*, const double*, double*, double*,  \
                             CGState*, RedBuf&, long, hipStream_t);                                               \
  template void build_block_inverse<K>(const MeshDev&, const double*, double*, hipStream_t);

ATMO_KRY_INST(1)
ATMO_KRY_INST(2)
ATMO_KRY_INST(3)

}  // namespace atmo
