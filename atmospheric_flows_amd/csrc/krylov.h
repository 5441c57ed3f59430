// Vector operations, deterministic reductions and device-resident Krylov
// state (krylov.hip).
//
// Reductions are two-level and order-fixed: a grid of RED_BLOCKS-limited
// blocks writes per-block partials, and the last block to finish (atomic
// ticket) sums them in block order and runs the finalization (alpha, beta,
// norms, convergence flag) on the device. Results therefore do not depend on
// scheduling, and Krylov scalars never travel to the host inside a solve
// except the convergence flag.
#pragma once
#include "common.h"
#include "ops.h"

namespace atmo {

constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 1024;
constexpr int MAX_GMRES = 64;

struct RedBuf {
  double* partials = nullptr;  // [RED_MAX_SLOTS][RED_MAX_BLOCKS]
  unsigned* counter = nullptr;
};
constexpr int RED_MAX_SLOTS = MAX_GMRES + 2;

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
// y = x * d (elementwise), optional scale
template <typename T>
void vec_mul(long n, const T* x, const T* d, T* y, hipStream_t s);

// out = x . y (sqrt when take_sqrt), written to device scalar *out
void vec_dot(long n, const double* x, const double* y, double* out, bool take_sqrt, RedBuf& rb, hipStream_t s);
// out = max |x - y| (y may be null), written to *out
void vec_linf(long n, const double* x, const double* y, double* out, RedBuf& rb, hipStream_t s);

// --- generic preconditioned CG pieces (host drives matvec/preconditioner)
enum DotMode { DOT_ALPHA = 0, DOT_GH_INIT = 1, DOT_BETA = 2, DOT_RES0 = 3 };
void cg_dot(long n, const double* x, const double* y, CGState* st, int mode, RedBuf& rb, hipStream_t s);
// x += alpha d ; g += alpha h ; res = |g| ; it++ ; convergence flag
void cg_update(long n, double* x, double* g, const double* d, const double* h, CGState* st, RedBuf& rb,
               hipStream_t s);
// d = -h + beta d
void cg_dir(long n, double* d, const double* h, const CGState* st, hipStream_t s);
// g = Ax - b (Ax in g) or g = -b (x == 0)
void cg_residual(long n, double* g, const double* b, bool have_ax, hipStream_t s);
// check after init: res0 = |g|, tol = tol_factor*tol_base (if tol_from_res0), done
void cg_init_check(long n, const double* g, CGState* st, bool tol_from_res0, RedBuf& rb, hipStream_t s);

// --- fused Jacobi-preconditioned CG on the A block (inner Schur solve)
// init: g = -b, d = 0, res0 = |b|, tol = tol_factor*res0, gh = g.P g, beta = 0
void jcg_init(long n, const double* b, const double* P, double* g, double* d, CGState* st, RedBuf& rb,
              hipStream_t s);
// d = -P g + beta d ; h = A d ; alpha = gh / (d.h)
template <int K>
void jcg_apply(const MeshDev& m, const double* rho, const double* P, const double* g, double* d, double* h,
               CGState* st, RedBuf& rb, long Ns, hipStream_t s);
// x += alpha d ; g += alpha h ; res, convergence ; beta = g.Pg / gh
void jcg_update(long n, double* x, double* g, const double* d, const double* h, const double* P, CGState* st,
                RedBuf& rb, hipStream_t s);

// --- GMRES pieces
// out[i] = w . V_i for i < nv (V column-major with leading dimension ldv)
void gm_mdot(long n, const double* w, const double* V, long ldv, int nv, double* out, RedBuf& rb, hipStream_t s);
// w -= sum_i coef[i] V_i  (coef device array)
void gm_maxpy(long n, double* w, const double* V, long ldv, int nv, const double* coef, hipStream_t s);
// x += sum_i y[i] V_i
void gm_update(long n, double* x, const double* V, long ldv, int nv, const double* y, hipStream_t s);
// y = x * scale_dev[0]^{-1}
void vec_scale_inv(long n, const double* x, const double* scal, double* y, hipStream_t s);

}  // namespace atmo
