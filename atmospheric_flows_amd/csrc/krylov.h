// Vector operations, deterministic reductions and device-resident Krylov
// state (krylov.hip).
//
// Reductions never use atomics: a producer kernel writes one partial per
// block (at most RED_MAX_BLOCKS blocks) and the consumer sums the partials in
// block order, either in a one-block finalize kernel (host-driven solvers)
// or redundantly in every block of the next kernel (the fused inner CG), so
// results do not depend on scheduling and no counter is contended across
// the chip's XCDs.
#pragma once
#include "common.h"
#include "ops.h"

namespace atmo {

constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 1024;
constexpr int MAX_GMRES = 64;
constexpr int RED_MAX_SLOTS = MAX_GMRES + 2;
// GMRES record of iteration j: w.V_0 .. w.V_{j+1}, then |w| after the
// projection in the last slot
constexpr int GM_REC = MAX_GMRES + 3;
// iterations of the Chebyshev eigenvalue estimate (their CG coefficients are
// recorded on the device)
constexpr int MAX_EIG_ITS = 256;

// Partial buffers: gen[RED_MAX_SLOTS][RED_MAX_BLOCKS] for host-driven
// reductions, pa[RED_MAX_BLOCKS] and pb[2][2][RED_MAX_BLOCKS] for the fused
// CG ping-pong.
struct RedBuf {
  double* gen = nullptr;
  double* pa = nullptr;
  double* pb = nullptr;
  static size_t doubles() { return (size_t)(RED_MAX_SLOTS + 1 + 4) * RED_MAX_BLOCKS; }
  void bind(double* base) {
    gen = base;
    pa = base + (size_t)RED_MAX_SLOTS * RED_MAX_BLOCKS;
    pb = pa + RED_MAX_BLOCKS;
  }
};

// State of one CG solve (deal.II SolverCG semantics), lives in device memory.
struct CGState {
  double gh, alpha, beta, res, tol, dh, tol_factor, res0;
  int it, done, max_its, pad;  // done: 0 iterate, 1 converged, -1 no convergence
};

int red_blocks(long n);

template <typename T>
void vec_axpby(long n, T a, const T* x, T b, T* y, hipStream_t s);  // y = a x + b y
template <typename T>
void vec_fill(long n, T v, T* y, hipStream_t s);
template <typename T>
void vec_copy(long n, const T* x, T* y, hipStream_t s);
template <typename TO, typename TI>
void vec_cast(long n, const TI* x, TO* y, hipStream_t s);
template <typename T>
void vec_mul(long n, const T* x, const T* d, T* y, hipStream_t s);  // y = x * d

// *out = x . y (sqrt when take_sqrt)
void vec_dot(long n, const double* x, const double* y, double* out, bool take_sqrt, RedBuf& rb, hipStream_t s);
// *out = max |x - y| (y may be null)
void vec_linf(long n, const double* x, const double* y, double* out, RedBuf& rb, hipStream_t s);

// --- generic preconditioned CG pieces (host drives matvec/preconditioner)
enum DotMode { DOT_ALPHA = 0, DOT_GH_INIT = 1, DOT_BETA = 2 };
// record (optional, device): receives alpha (DOT_ALPHA) or beta (DOT_BETA)
// of this iteration when the solve has not finished
void cg_dot(long n, const double* x, const double* y, CGState* st, int mode, RedBuf& rb, hipStream_t s,
            double* record = nullptr);
// x += alpha d ; g += alpha h ; res = |g| ; it++ ; convergence flag
void cg_update(long n, double* x, double* g, const double* d, const double* h, CGState* st, RedBuf& rb,
               hipStream_t s);
// d = -h + beta d
void cg_dir(long n, double* d, const double* h, const CGState* st, hipStream_t s);
// g = Ax - b (Ax already in g) or g = -b
void cg_residual(long n, double* g, const double* b, bool have_ax, hipStream_t s);
// res0 = |g|, tol = tol_factor*res0 when tol_from_res0, it = 0, check
void cg_init_check(long n, const double* g, CGState* st, bool tol_from_res0, RedBuf& rb, hipStream_t s);

// --- fused CG on the velocity block A (inner Schur solve), zero initial
// guess, tolerance tol_factor*|b|, Jacobi preconditioner (P = inverse
// diagonal). Iteration i (1-based, passed by the host) is jcg_apply(i) then
// jcg_update(i).
void jcg_init(long Ns, const MeshDev& m, int K, const double* b, const double* P, double* g, double* d, CGState* st,
              RedBuf& rb, hipStream_t s);
template <int K>
void jcg_apply(int i, const MeshDev& m, const double* rho, const double* P, const double* g, double* d, double* h,
               CGState* st, RedBuf& rb, long Ns, hipStream_t s);
void jcg_update(int i, long Ns, const MeshDev& m, int K, double* x, double* g, const double* d, const double* h,
                const double* P, CGState* st, RedBuf& rb, hipStream_t s);

// FP64 inverse of every cell block of A (rho-weighted mass), [cell][i][j]
// (symmetric, so row i is also column i)
template <int K>
void build_block_inverse(const MeshDev& m, const double* rho, double* Ainv, hipStream_t s);

// --- GMRES pieces
// out[i] = w . V_i for i < nv (V column-major with leading dimension ldv)
void gm_mdot(long n, const double* w, const double* V, long ldv, int nv, double* out, RedBuf& rb, hipStream_t s);
// w -= sum_i sig[i]^2 dots[i] V_i (device arrays) and *norm = |w| of the
// result, *inv_norm = 1/|w| (0 for w = 0) (device scalars)
void gm_maxpy_norm(long n, double* w, const double* V, long ldv, int nv, const double* dots, const double* sig,
                   double* norm, double* inv_norm, RedBuf& rb, hipStream_t s);
// out[s] = sum_{b < G} part[s*pstride + b] for s < nslot (fixed order)
void gm_sum_slots(const double* part, long pstride, int G, int nslot, double* out, hipStream_t s);
// *norm = |v|, *inv_norm = 1/|v| (0 for v = 0) (device scalars)
void gm_norm(long n, const double* v, double* norm, double* inv_norm, RedBuf& rb, hipStream_t s);
// x += sum_i y[i] V_i
void gm_update(long n, double* x, const double* V, long ldv, int nv, const double* y, hipStream_t s);

}  // namespace atmo
