// C ABI of libatmo_hip.so (include/atmo_hip.h): argument checking, exception
// to status mapping and library-owned device vectors.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/atmo_hip.h"
#include "diag.h"
#include "engine.h"
#include "ensemble.h"
#include "krylov.h"
#include "vtu.h"

using namespace atmo;

struct atmo_ctx {
  std::unique_ptr<Engine> eng;
  MeshDesc mesh;
  int degree = 0;
  std::string err;
};

struct atmo_ensemble {
  std::unique_ptr<Ensemble> ens;
};

struct atmo_vec {
  double* p = nullptr;
  size_t n = 0;
  int block = 0;
};

static std::mutex g_err_mu;
static std::string g_err;

static void set_err(atmo_ctx* c, const std::string& s) {
  if (c) c->err = s;
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = s;
}

template <class F>
static atmo_status guarded(atmo_ctx* c, F f) {
  try {
    return f();
  } catch (const std::invalid_argument& e) {
    set_err(c, e.what());
    return ATMO_INVALID_ARG;
  } catch (const std::bad_alloc& e) {
    set_err(c, e.what());
    return ATMO_OOM;
  } catch (const DeviceError& e) {
    set_err(c, e.what());
    return std::string(e.what()).find("out of memory") != std::string::npos ? ATMO_OOM : ATMO_DEVICE;
  } catch (const std::runtime_error& e) {
    set_err(c, e.what());
    const std::string w = e.what();
    if (w.find("no convergence") != std::string::npos) return ATMO_NO_CONVERGENCE;
    if (w.find("non-finite") != std::string::npos) return ATMO_NONFINITE;
    return ATMO_DEVICE;
  } catch (...) {
    set_err(c, "unknown error");
    return ATMO_DEVICE;
  }
}

// the same for context-free entry points defined in other files (errors read with atmo_last_error(NULL))
atmo_status atmo_guarded_noctx(const std::function<atmo_status()>& f) { return guarded(nullptr, f); }

static EngineParams to_engine(const atmo_params& p) {
  EngineParams e;
  e.Ma = p.Ma;
  e.Fr = p.Fr;
  e.dt = p.dt;
  e.gamma_gas = p.gamma_gas;
  e.grav[0] = p.grav_x;
  e.grav[1] = p.grav_z;
  e.grav_factor = p.grav_factor;
  e.eps = p.eps;
  e.max_its = p.max_iterations;
  e.schur_inner_tol = p.schur_inner_tol;
  e.schur_inner_max = p.schur_inner_max;
  e.a_solve_tol = p.a_solve_tol;
  e.fp_max_iter = p.fp_max_iter;
  e.fp_tol = p.fp_tol;
  e.fp_lock = p.fp_lock;
  e.mg_min_cells = p.mg_min_cells;
  e.cheb_degree = p.cheb_degree;
  e.cheb_range = p.cheb_range;
  e.cheb_eig_its = p.cheb_eig_its;
  e.gmres_restart = p.gmres_restart;
  e.precision = p.precision;
  e.mg_matrix_cores = p.mg_matrix_cores;
  e.arnoldi_waves = p.arnoldi_waves_per_simd;
  return e;
}

static void check_params(const atmo_params& p) {
  if (!(p.dt > 0) || !(p.Ma > 0) || !(p.Fr > 0) || !(p.gamma_gas > 1))
    throw std::invalid_argument("params: dt, Ma, Fr must be > 0 and gamma > 1");
  if (p.max_iterations < 1 || p.schur_inner_max < 1 || p.fp_max_iter < 1 || p.cheb_degree < 1 ||
      p.cheb_eig_its < 1)
    throw std::invalid_argument("params: iteration limits must be >= 1");
  if (p.cheb_eig_its > MAX_EIG_ITS) throw std::invalid_argument("params: cheb_eig_its must be <= 256");
  if (p.gmres_restart < 1 || p.gmres_restart > MAX_GMRES - 2)
    throw std::invalid_argument("params: gmres_restart must be in [1, 62]");
  if (p.precision < 0 || p.precision > 2) throw std::invalid_argument("params: precision must be 0, 1 or 2");
  if (p.mg_matrix_cores != 0 && p.mg_matrix_cores != 1)
    throw std::invalid_argument("params: mg_matrix_cores must be 0 or 1");
  if (p.arnoldi_waves_per_simd != 0 && p.arnoldi_waves_per_simd != 3 && p.arnoldi_waves_per_simd != 4)
    throw std::invalid_argument("params: arnoldi_waves_per_simd must be 0, 3 or 4");
  if (!(p.eps >= 0) || !(p.schur_inner_tol >= 0) || !(p.a_solve_tol >= 0))
    throw std::invalid_argument("params: tolerances must be >= 0");
  if (p.mg_min_cells < 1) throw std::invalid_argument("params: mg_min_cells must be >= 1");
}

// components of a block's field: the velocity has dim of them
static size_t block_size(const atmo_ctx* c, int block) {
  return (size_t)(block == ATMO_BLOCK_VELOCITY ? c->eng->dim() : 1) * (size_t)c->eng->ns();
}

static void check_vec(const atmo_ctx* c, const atmo_vec* v, int block) {
  if (!v) throw std::invalid_argument("null vector");
  const size_t want = block_size(c, block);
  if (v->n != want) throw std::invalid_argument("vector does not belong to the expected block");
}

extern "C" {

void atmo_params_default(atmo_params* p) {
  if (!p) return;
  EngineParams e;
  p->Ma = e.Ma;
  p->Fr = e.Fr;
  p->dt = e.dt;
  p->gamma_gas = e.gamma_gas;
  p->grav_x = e.grav[0];
  p->grav_z = e.grav[1];
  p->grav_factor = e.grav_factor;
  p->eps = e.eps;
  p->max_iterations = e.max_its;
  p->schur_inner_tol = e.schur_inner_tol;
  p->schur_inner_max = e.schur_inner_max;
  p->a_solve_tol = e.a_solve_tol;
  p->fp_max_iter = e.fp_max_iter;
  p->fp_tol = e.fp_tol;
  p->fp_lock = e.fp_lock;
  p->mg_min_cells = e.mg_min_cells;
  p->cheb_degree = e.cheb_degree;
  p->cheb_range = e.cheb_range;
  p->cheb_eig_its = e.cheb_eig_its;
  p->gmres_restart = e.gmres_restart;
  p->precision = e.precision;
  p->mg_matrix_cores = e.mg_matrix_cores;
  p->arnoldi_waves_per_simd = e.arnoldi_waves;
}

const char* atmo_version(void) { return "atmo_hip 0.1 (gfx950)"; }

const char* atmo_last_error(const atmo_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  std::lock_guard<std::mutex> lk(g_err_mu);
  return g_err.c_str();
}

// Validated copy of a mesh description.
static MeshDesc to_mesh(const atmo_mesh_desc* mesh) {
  if (mesh->degree < 1 || mesh->degree > 3) throw std::invalid_argument("degree must be 1, 2 or 3");
  if (mesh->nx < 1 || mesh->nz < 1 || !mesh->h) throw std::invalid_argument("mesh: nx, nz >= 1 and h required");
  if (mesh->periodic && mesh->nx < 2) throw std::invalid_argument("mesh: periodic mesh needs nx >= 2");
  if (!(mesh->Lx > 0) || !(mesh->H > 0)) throw std::invalid_argument("mesh: Lx, H must be > 0");
  for (int i = 0; i <= mesh->nx; ++i)
    if (!std::isfinite(mesh->h[i]) || mesh->h[i] >= mesh->H) throw std::invalid_argument("mesh: bad terrain height");
  // x-periodic: the seam's two sides are one face, so its end columns must coincide (deal.II's
  // collect_periodic_faces rejects unmatched faces); otherwise the seam flux is not conservative
  if (mesh->periodic && std::fabs(mesh->h[0] - mesh->h[mesh->nx]) > 1e-12 * mesh->H)
    throw std::invalid_argument("mesh: periodic mesh needs h[0] == h[nx]");
  MeshDesc m;
  m.nx = mesh->nx;
  m.nz = mesh->nz;
  m.periodic = mesh->periodic ? 1 : 0;
  m.Lx = mesh->Lx;
  m.H = mesh->H;
  m.h.assign(mesh->h, mesh->h + mesh->nx + 1);
  return m;
}

atmo_status atmo_ctx_create(const atmo_mesh_desc* mesh, const atmo_params* params, atmo_ctx** out) {
  return guarded(nullptr, [&]() {
    if (!mesh || !params || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    MeshDesc m = to_mesh(mesh);
    check_params(*params);
    auto c = std::make_unique<atmo_ctx>();
    c->mesh = std::move(m);
    c->degree = mesh->degree;
    c->eng = Engine::create(mesh->degree, c->mesh, to_engine(*params));
    *out = c.release();
    return ATMO_OK;
  });
}

atmo_status atmo_ctx_create_amr(const atmo_mesh_desc* base, const int* leaves, int nleaves, const atmo_params* params,
                                atmo_ctx** out) {
  return guarded(nullptr, [&]() {
    if (!base || !leaves || !params || !out) throw std::invalid_argument("null argument");
    if (nleaves < 1) throw std::invalid_argument("locally refined mesh: no leaves");
    *out = nullptr;
    MeshDesc m = to_mesh(base);
    check_params(*params);
    std::vector<QCell> cells(nleaves);
    for (int c = 0; c < nleaves; ++c)
      cells[c] = {leaves[5 * c], leaves[5 * c + 1], leaves[5 * c + 2], leaves[5 * c + 3], leaves[5 * c + 4]};
    auto c = std::make_unique<atmo_ctx>();
    c->mesh = std::move(m);
    c->degree = base->degree;
    c->eng = Engine::create_amr(base->degree, c->mesh, cells, to_engine(*params));
    *out = c.release();
    return ATMO_OK;
  });
}

atmo_status atmo_ctx_init_hill_state(atmo_ctx* ctx, double theta0, double p0, double N, double U, double g, double R,
                                     double gamma_gas, double perturbation, uint64_t seed, double* rho, double* u,
                                     double* p) {
  return guarded(ctx, [&]() {
    if (!ctx || !rho || !u || !p) throw std::invalid_argument("null argument");
    if (ctx->eng->shell()) throw std::invalid_argument("hill state: not on a shell context");
    const std::vector<double>& geo = ctx->eng->cell_geometry();
    if (geo.empty())
      hill_initial_state(ctx->degree, ctx->mesh, theta0, p0, N, U, g, R, gamma_gas, perturbation, seed, rho, u, p);
    else
      hill_initial_state_cells(ctx->degree, geo, theta0, p0, N, U, g, R, gamma_gas, perturbation, seed, rho, u, p);
    return ATMO_OK;
  });
}

atmo_status atmo_ctx_destroy(atmo_ctx* ctx) {
  if (!ctx) return ATMO_INVALID_ARG;
  delete ctx;
  return ATMO_OK;
}

atmo_status atmo_set_params(atmo_ctx* ctx, const atmo_params* params) {
  return guarded(ctx, [&]() {
    if (!ctx || !params) throw std::invalid_argument("null argument");
    check_params(*params);
    ctx->eng->set_params(to_engine(*params));
    return ATMO_OK;
  });
}

atmo_status atmo_ctx_info(atmo_ctx* ctx, int* ncells, long* scalar_dofs, int* nlevels) {
  if (!ctx) return ATMO_INVALID_ARG;
  if (ncells) *ncells = ctx->eng->ncells();
  if (scalar_dofs) *scalar_dofs = ctx->eng->ns();
  if (nlevels) *nlevels = ctx->eng->nlevels();
  return ATMO_OK;
}

atmo_status atmo_sync(atmo_ctx* ctx) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    ATMO_HIP_CHECK(hipStreamSynchronize(ctx->eng->stream()));
    return ATMO_OK;
  });
}

atmo_status atmo_get_stats(atmo_ctx* ctx, atmo_stats* st) {
  if (!ctx || !st) return ATMO_INVALID_ARG;
  const EngineStats& s = ctx->eng->stats();
  st->cg_its = s.cg_its;
  st->gmres_its = s.gmres_its;
  st->inner_cg_its = s.inner_cg_its;
  st->fp_its = s.fp_its;
  st->mg_cycles = s.mg_cycles;
  st->coarse_its = s.coarse_its;
  st->schur_matvecs = s.schur_matvecs;
  st->host_syncs = s.host_syncs;
  return ATMO_OK;
}

atmo_status atmo_reset_stats(atmo_ctx* ctx) {
  if (!ctx) return ATMO_INVALID_ARG;
  ctx->eng->stats() = EngineStats();
  return ATMO_OK;
}

atmo_status atmo_vec_alloc(atmo_ctx* ctx, int block, atmo_vec** out) {
  return guarded(ctx, [&]() {
    if (!ctx || !out) throw std::invalid_argument("null argument");
    if (block < 0 || block > 2) throw std::invalid_argument("block must be 0 (velocity), 1 (pressure), 2 (density)");
    auto v = std::make_unique<atmo_vec>();
    v->block = block;
    v->n = block_size(ctx, block);
    ATMO_HIP_CHECK(hipMalloc(&v->p, v->n * sizeof(double)));
    ATMO_HIP_CHECK(hipMemsetAsync(v->p, 0, v->n * sizeof(double), ctx->eng->stream()));
    *out = v.release();
    return ATMO_OK;
  });
}

atmo_status atmo_vec_free(atmo_ctx* ctx, atmo_vec* v) {
  if (!v) return ATMO_INVALID_ARG;
  if (ctx) (void)hipStreamSynchronize(ctx->eng->stream());
  if (v->p) (void)hipFree(v->p);
  delete v;
  return ATMO_OK;
}

atmo_status atmo_vec_size(const atmo_vec* v, size_t* n) {
  if (!v || !n) return ATMO_INVALID_ARG;
  *n = v->n;
  return ATMO_OK;
}

atmo_status atmo_vec_upload(atmo_ctx* ctx, atmo_vec* v, const double* host, size_t n) {
  return guarded(ctx, [&]() {
    if (!ctx || !v || !host) throw std::invalid_argument("null argument");
    if (n != v->n) throw std::invalid_argument("upload: size mismatch");
    ATMO_HIP_CHECK(hipMemcpyAsync(v->p, host, n * sizeof(double), hipMemcpyHostToDevice, ctx->eng->stream()));
    ATMO_HIP_CHECK(hipStreamSynchronize(ctx->eng->stream()));
    return ATMO_OK;
  });
}

atmo_status atmo_vec_download(atmo_ctx* ctx, const atmo_vec* v, double* host, size_t n) {
  return guarded(ctx, [&]() {
    if (!ctx || !v || !host) throw std::invalid_argument("null argument");
    if (n != v->n) throw std::invalid_argument("download: size mismatch");
    ATMO_HIP_CHECK(hipMemcpyAsync(host, v->p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->eng->stream()));
    ATMO_HIP_CHECK(hipStreamSynchronize(ctx->eng->stream()));
    return ATMO_OK;
  });
}

atmo_status atmo_vec_device_ptr(atmo_vec* v, double** dev) {
  if (!v || !dev) return ATMO_INVALID_ARG;
  *dev = v->p;
  return ATMO_OK;
}

atmo_status atmo_vec_axpby(atmo_ctx* ctx, double a, const atmo_vec* x, double b, atmo_vec* y) {
  return guarded(ctx, [&]() {
    if (!ctx || !x || !y) throw std::invalid_argument("null argument");
    if (x->n != y->n) throw std::invalid_argument("axpby: vectors of different blocks");
    ctx->eng->axpby((long)y->n, a, x->p, b, y->p);
    return ATMO_OK;
  });
}

atmo_status atmo_vec_dot(atmo_ctx* ctx, const atmo_vec* x, const atmo_vec* y, double* result) {
  return guarded(ctx, [&]() {
    if (!ctx || !x || !y || !result) throw std::invalid_argument("null argument");
    if (x->n != y->n) throw std::invalid_argument("dot: vectors of different blocks");
    *result = ctx->eng->dot((long)x->n, x->p, y->p);
    return ATMO_OK;
  });
}

atmo_status atmo_vec_linf(atmo_ctx* ctx, const atmo_vec* x, double* result) {
  return guarded(ctx, [&]() {
    if (!ctx || !x || !result) throw std::invalid_argument("null argument");
    *result = ctx->eng->linf((long)x->n, x->p);
    return ATMO_OK;
  });
}

atmo_status atmo_set_stages(atmo_ctx* ctx, int hyperbolic_stage, int ns_stage) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    ctx->eng->set_stages(hyperbolic_stage, ns_stage);
    return ATMO_OK;
  });
}

atmo_status atmo_set_rho_for_fixed(atmo_ctx* ctx, const atmo_vec* rho) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    check_vec(ctx, rho, ATMO_BLOCK_DENSITY);
    ctx->eng->set_rho_for_fixed(rho->p);
    return ATMO_OK;
  });
}

atmo_status atmo_set_pres_fixed(atmo_ctx* ctx, const atmo_vec* p) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    check_vec(ctx, p, ATMO_BLOCK_PRESSURE);
    ctx->eng->set_pres_fixed(p->p);
    return ATMO_OK;
  });
}

// positional source lists: rho/p entries are scalar, u entries are velocity
static std::vector<const double*> gather(const atmo_ctx* c, const atmo_vec* const* src, int nsrc) {
  if (nsrc < 1 || nsrc > 11 || !src) throw std::invalid_argument("bad source list");
  std::vector<const double*> out(nsrc);
  for (int i = 0; i < nsrc; ++i) {
    if (!src[i]) throw std::invalid_argument("null source vector");
    if (src[i]->n != block_size(c, src[i]->block))
      throw std::invalid_argument("source vector of another context");
    out[i] = src[i]->p;
  }
  return out;
}

// expected block of position i for the residual source lists
static void check_positions(const atmo_vec* const* src, int nsrc, const int* blocks) {
  for (int i = 0; i < nsrc; ++i)
    if (src[i]->block != blocks[i]) throw std::invalid_argument("source list: wrong block at position " + std::to_string(i));
}

atmo_status atmo_rhs_rho(atmo_ctx* ctx, atmo_vec* dst, const atmo_vec* const* src, int nsrc) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    check_vec(ctx, dst, ATMO_BLOCK_DENSITY);
    auto s = gather(ctx, src, nsrc);
    const int bl[6] = {2, 0, 2, 0, 2, 0};
    if (nsrc > 6) throw std::invalid_argument("rhs_rho: at most 6 sources");
    check_positions(src, nsrc, bl);
    ctx->eng->rhs_rho(s.data(), nsrc, dst->p);
    return ATMO_OK;
  });
}

atmo_status atmo_rhs_pressure(atmo_ctx* ctx, atmo_vec* dst, const atmo_vec* const* src, int nsrc) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    check_vec(ctx, dst, ATMO_BLOCK_PRESSURE);
    auto s = gather(ctx, src, nsrc);
    const int bl[11] = {2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0};
    // stage 1: {rho,u,p,rho2,u_fix}; stage 2: {...,rho3,u_fix}
    if (nsrc == 5) {
      const int b1[5] = {2, 0, 1, 2, 0};
      check_positions(src, nsrc, b1);
    } else if (nsrc == 8) {
      const int b2[8] = {2, 0, 1, 2, 0, 1, 2, 0};
      check_positions(src, nsrc, b2);
    } else {
      check_positions(src, nsrc, bl);
    }
    ctx->eng->rhs_pressure(s.data(), nsrc, dst->p);
    return ATMO_OK;
  });
}

atmo_status atmo_rhs_velocity(atmo_ctx* ctx, atmo_vec* dst, const atmo_vec* const* src, int nsrc) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    check_vec(ctx, dst, ATMO_BLOCK_VELOCITY);
    auto s = gather(ctx, src, nsrc);
    const int bl[9] = {2, 0, 1, 2, 0, 1, 2, 0, 1};
    if (nsrc > 9) throw std::invalid_argument("rhs_velocity: at most 9 sources");
    check_positions(src, nsrc, bl);
    ctx->eng->rhs_velocity(s.data(), nsrc, dst->p);
    return ATMO_OK;
  });
}

atmo_status atmo_vmult_pressure(atmo_ctx* ctx, atmo_vec* dst_u, const atmo_vec* src_p) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    check_vec(ctx, dst_u, ATMO_BLOCK_VELOCITY);
    check_vec(ctx, src_p, ATMO_BLOCK_PRESSURE);
    ctx->eng->vmult_pressure(src_p->p, dst_u->p);
    return ATMO_OK;
  });
}

atmo_status atmo_vmult_enthalpy(atmo_ctx* ctx, atmo_vec* dst_p, const atmo_vec* src_u) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    check_vec(ctx, dst_p, ATMO_BLOCK_PRESSURE);
    check_vec(ctx, src_u, ATMO_BLOCK_VELOCITY);
    ctx->eng->vmult_enthalpy(src_u->p, dst_p->p);
    return ATMO_OK;
  });
}

atmo_status atmo_vmult(atmo_ctx* ctx, atmo_vec* dst, const atmo_vec* src) {
  return guarded(ctx, [&]() {
    if (!ctx || !dst || !src) throw std::invalid_argument("null argument");
    if (dst->n != src->n) throw std::invalid_argument("vmult: size mismatch");
    ctx->eng->vmult(src->p, dst->p);
    ATMO_HIP_CHECK(hipStreamSynchronize(ctx->eng->stream()));
    return ATMO_OK;
  });
}

atmo_status atmo_inverse_diagonal(atmo_ctx* ctx, atmo_vec* dst) {
  return guarded(ctx, [&]() {
    if (!ctx || !dst) throw std::invalid_argument("null argument");
    ctx->eng->inverse_diagonal(dst->p);
    return ATMO_OK;
  });
}

atmo_status atmo_solve_cg_mg(atmo_ctx* ctx, int ns_stage, atmo_vec* x, const atmo_vec* b, double tol_factor,
                             int* iterations) {
  return guarded(ctx, [&]() {
    if (!ctx || !x || !b) throw std::invalid_argument("null argument");
    if (ns_stage < 1 || ns_stage > 6 || ns_stage == 2) throw std::invalid_argument("ns_stage must be 1, 3, 4, 5 or 6");
    const int blk = (ns_stage == 3 || ns_stage == 5) ? ATMO_BLOCK_VELOCITY : (ns_stage == 6 ? ATMO_BLOCK_PRESSURE : ATMO_BLOCK_DENSITY);
    check_vec(ctx, x, blk);
    check_vec(ctx, b, blk);
    const int its = ctx->eng->solve_mass_mg(ns_stage, x->p, b->p, tol_factor);
    if (iterations) *iterations = its;
    return ATMO_OK;
  });
}

atmo_status atmo_solve_schur_gmres(atmo_ctx* ctx, atmo_vec* x, const atmo_vec* b, double tol_factor,
                                   int* iterations) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    check_vec(ctx, x, ATMO_BLOCK_PRESSURE);
    check_vec(ctx, b, ATMO_BLOCK_PRESSURE);
    const int its = ctx->eng->solve_schur(x->p, b->p, tol_factor);
    if (iterations) *iterations = its;
    return ATMO_OK;
  });
}

atmo_status atmo_state_upload(atmo_ctx* ctx, const double* rho, const double* u, const double* p) {
  return guarded(ctx, [&]() {
    if (!ctx || !rho || !u || !p) throw std::invalid_argument("null argument");
    const long Ns = ctx->eng->ns();
    const int nd = ctx->eng->dim();
    for (long i = 0; i < Ns; ++i) {
      bool ok = std::isfinite(rho[i]) && std::isfinite(p[i]) && rho[i] > 0;
      for (int c = 0; c < nd; ++c) ok = ok && std::isfinite(u[c * Ns + i]);
      if (!ok) throw std::runtime_error("state_upload: non-finite state or non-positive density");
    }
    hipStream_t s = ctx->eng->stream();
    ATMO_HIP_CHECK(hipMemcpyAsync(ctx->eng->state_ptr(0), rho, Ns * sizeof(double), hipMemcpyHostToDevice, s));
    ATMO_HIP_CHECK(hipMemcpyAsync(ctx->eng->state_ptr(1), u, nd * Ns * sizeof(double), hipMemcpyHostToDevice, s));
    ATMO_HIP_CHECK(hipMemcpyAsync(ctx->eng->state_ptr(2), p, Ns * sizeof(double), hipMemcpyHostToDevice, s));
    ATMO_HIP_CHECK(hipStreamSynchronize(s));
    ctx->eng->load_state(ctx->eng->state_ptr(0), ctx->eng->state_ptr(1), ctx->eng->state_ptr(2));
    return ATMO_OK;
  });
}

atmo_status atmo_state_download(atmo_ctx* ctx, double* rho, double* u, double* p) {
  return guarded(ctx, [&]() {
    if (!ctx || !rho || !u || !p) throw std::invalid_argument("null argument");
    const long Ns = ctx->eng->ns();
    hipStream_t s = ctx->eng->stream();
    ATMO_HIP_CHECK(hipMemcpyAsync(rho, ctx->eng->state_ptr(0), Ns * sizeof(double), hipMemcpyDeviceToHost, s));
    ATMO_HIP_CHECK(
        hipMemcpyAsync(u, ctx->eng->state_ptr(1), ctx->eng->dim() * Ns * sizeof(double), hipMemcpyDeviceToHost, s));
    ATMO_HIP_CHECK(hipMemcpyAsync(p, ctx->eng->state_ptr(2), Ns * sizeof(double), hipMemcpyDeviceToHost, s));
    ATMO_HIP_CHECK(hipStreamSynchronize(s));
    return ATMO_OK;
  });
}

atmo_status atmo_imex_step(atmo_ctx* ctx, int nsteps) {
  return guarded(ctx, [&]() {
    if (!ctx) throw std::invalid_argument("null context");
    if (nsteps < 0) throw std::invalid_argument("nsteps must be >= 0");
    for (int n = 0; n < nsteps; ++n) ctx->eng->step();
    ATMO_HIP_CHECK(hipStreamSynchronize(ctx->eng->stream()));
    return ATMO_OK;
  });
}

atmo_status atmo_step_diagnostics(atmo_ctx* ctx, int of_state, atmo_step_diag* out) {
  return guarded(ctx, [&]() {
    if (!ctx || !out) throw std::invalid_argument("null argument");
    double rec[DG_SLOTS];
    ctx->eng->step_diagnostics(rec, of_state != 0);
    for (int s = 0; s < 3; ++s) {
      out->min_density[s] = rec[DG_MIN_RHO + s];
      out->max_density[s] = rec[DG_MAX_RHO + s];
    }
    out->max_celerity = rec[DG_MAX_CEL];
    out->max_velocity = rec[DG_MAX_VEL];
    double bad = 0.0;
    for (int s = 0; s < 4; ++s) bad += rec[DG_NONFINITE + s];
    out->nonfinite = (long)bad;
    const MeshDesc& m = ctx->mesh;
    out->min_cell_diameter = ctx->eng->min_cell_diameter() > 0 ? ctx->eng->min_cell_diameter()
                                                                 : min_cell_diameter(m.nx, m.nz, m.Lx, m.H, m.h.data());
    // Courant numbers as printed by EulerSolver::run (Atmospheric_Flow.cc:1225-1235)
    const EngineParams& p = ctx->eng->params();
    const double f = p.dt * ctx->degree * std::sqrt((double)ctx->eng->dim()) / out->min_cell_diameter;
    out->cfl_c = f * out->max_celerity / p.Ma;
    out->cfl_u = f * out->max_velocity;
    return ATMO_OK;
  });
}

atmo_status atmo_init_hill_state(const atmo_mesh_desc* mesh, double theta0, double p0, double N, double U, double g,
                                 double R, double gamma_gas, double perturbation, uint64_t seed, double* rho,
                                 double* u, double* p) {
  return guarded(nullptr, [&]() {
    if (!mesh || !mesh->h || !rho || !u || !p) throw std::invalid_argument("null argument");
    const MeshDesc m = to_mesh(mesh);
    if (!(N > 0) || !(theta0 > 0) || !(p0 > 0) || !(g > 0)) throw std::invalid_argument("bad background parameters");
    hill_initial_state(mesh->degree, m, theta0, p0, N, U, g, R, gamma_gas, perturbation, seed, rho, u, p);
    return ATMO_OK;
  });
}

atmo_status atmo_write_vtu(atmo_ctx* ctx, const char* path, int high_order) {
  return guarded(ctx, [&]() {
    if (!ctx || !path) throw std::invalid_argument("null argument");
    if (ctx->eng->shell()) throw std::invalid_argument("write_vtu: not available on a shell context");
    const long Ns = ctx->eng->ns();
    std::vector<double> rho(Ns), u(2 * Ns), p(Ns);
    hipStream_t s = ctx->eng->stream();
    ATMO_HIP_CHECK(hipMemcpyAsync(rho.data(), ctx->eng->state_ptr(0), Ns * sizeof(double), hipMemcpyDeviceToHost, s));
    ATMO_HIP_CHECK(hipMemcpyAsync(u.data(), ctx->eng->state_ptr(1), 2 * Ns * sizeof(double), hipMemcpyDeviceToHost, s));
    ATMO_HIP_CHECK(hipMemcpyAsync(p.data(), ctx->eng->state_ptr(2), Ns * sizeof(double), hipMemcpyDeviceToHost, s));
    ATMO_HIP_CHECK(hipStreamSynchronize(s));
    write_vtu(ctx->degree, ctx->mesh, ctx->eng->cell_geometry(), rho.data(), u.data(), p.data(), path, high_order != 0);
    return ATMO_OK;
  });
}

atmo_status atmo_write_vtu_host(const atmo_mesh_desc* mesh, const double* rho, const double* u, const double* p,
                                const char* path, int high_order) {
  return guarded(nullptr, [&]() {
    if (!mesh || !mesh->h || !rho || !u || !p || !path) throw std::invalid_argument("null argument");
    if (mesh->degree < 1 || mesh->degree > 3 || mesh->nx < 1 || mesh->nz < 1)
      throw std::invalid_argument("bad mesh description");
    write_vtu(mesh->degree, to_mesh(mesh), {}, rho, u, p, path, high_order != 0);
    return ATMO_OK;
  });
}

atmo_status atmo_time_kernel(atmo_ctx* ctx, int which, int nrep, double* ms_per_launch, double* bytes_per_launch) {
  return guarded(ctx, [&]() {
    if (!ctx || !ms_per_launch || !bytes_per_launch || nrep < 1) throw std::invalid_argument("bad argument");
    ctx->eng->time_kernel(which, nrep, ms_per_launch, bytes_per_launch);
    return ATMO_OK;
  });
}

static void to_tuning(const KernelTuning& kt, atmo_kernel_tuning* o) {
  for (int i = 0; i < 2; ++i) {
    o->arnoldi_us[i] = kt.arnoldi_us[i];
    o->sweep_us[i] = kt.sweep_us[i];
  }
  o->arnoldi_waves_per_simd = kt.arnoldi_waves;
  o->mg_matrix_cores = kt.mg_matrix_cores;
}

atmo_status atmo_tune_kernels(atmo_ctx* ctx, int nrep, atmo_kernel_tuning* out) {
  return guarded(ctx, [&]() {
    if (!ctx || !out || nrep < 1) throw std::invalid_argument("bad argument");
    to_tuning(ctx->eng->tune_kernels(nrep), out);
    return ATMO_OK;
  });
}


// ------------------------------------------------------------ spherical shell
static ShellDesc to_shell(const atmo_shell_desc* d) {
  ShellDesc s;
  s.refinements = d->refinements;
  s.panels = d->panels;
  s.r_inner = d->r_inner;
  s.r_outer = d->r_outer;
  s.planet = d->planet;
  if (s.panels != 6 && s.panels != 12) throw std::invalid_argument("shell: panels must be 6 or 12");
  if (s.refinements < 0 || s.refinements > 7) throw std::invalid_argument("shell: refinements must be in [0, 7]");
  if (!(s.r_inner > 0) || !(s.r_outer > s.r_inner) || !std::isfinite(s.r_outer) || !(s.planet > 0) ||
      !std::isfinite(s.planet))
    throw std::invalid_argument("shell: need 0 < r_inner < r_outer and planet > 0");
  return s;
}

static const ShellMesh& shell_of(atmo_ctx* ctx) {
  if (!ctx->eng->shell()) throw std::invalid_argument("not a shell context");
  return *ctx->eng->shell();
}

atmo_status atmo_ctx_create_shell(const atmo_shell_desc* shell, const atmo_params* params, atmo_ctx** out) {
  return guarded(nullptr, [&]() {
    if (!shell || !params || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    const ShellDesc d = to_shell(shell);
    check_params(*params);
    auto c = std::make_unique<atmo_ctx>();
    c->degree = 1;
    c->eng = Engine::create_shell(d, to_engine(*params));
    *out = c.release();
    return ATMO_OK;
  });
}

atmo_status atmo_ctx_dim(atmo_ctx* ctx, int* dim) {
  if (!ctx || !dim) return ATMO_INVALID_ARG;
  *dim = ctx->eng->dim();
  return ATMO_OK;
}

atmo_status atmo_shell_dof_points(atmo_ctx* ctx, double* xyz) {
  return guarded(ctx, [&]() {
    if (!ctx || !xyz) throw std::invalid_argument("null argument");
    shell_dof_points(shell_of(ctx), xyz);
    return ATMO_OK;
  });
}

atmo_status atmo_shell_sbr_state(atmo_ctx* ctx, double Omega, double a, double g, double RT, double* rho, double* u,
                                 double* p) {
  return guarded(ctx, [&]() {
    if (!ctx || !rho || !u || !p) throw std::invalid_argument("null argument");
    if (!(RT > 0) || !(a > 0) || !std::isfinite(Omega) || !std::isfinite(g))
      throw std::invalid_argument("sbr state: need RT > 0, a > 0");
    shell_sbr_state(shell_of(ctx), SbrParams{Omega, a, g, RT}, rho, u, p);
    return ATMO_OK;
  });
}

atmo_status atmo_shell_l2_errors(atmo_ctx* ctx, double Omega, double a, double g, double RT, double* out) {
  return guarded(ctx, [&]() {
    if (!ctx || !out) throw std::invalid_argument("null argument");
    const ShellMesh& m = shell_of(ctx);
    const long Ns = ctx->eng->ns();
    std::vector<double> rho(Ns), u(3 * Ns), p(Ns);
    hipStream_t s = ctx->eng->stream();
    ATMO_HIP_CHECK(hipMemcpyAsync(rho.data(), ctx->eng->state_ptr(0), Ns * sizeof(double), hipMemcpyDeviceToHost, s));
    ATMO_HIP_CHECK(hipMemcpyAsync(u.data(), ctx->eng->state_ptr(1), 3 * Ns * sizeof(double), hipMemcpyDeviceToHost, s));
    ATMO_HIP_CHECK(hipMemcpyAsync(p.data(), ctx->eng->state_ptr(2), Ns * sizeof(double), hipMemcpyDeviceToHost, s));
    ATMO_HIP_CHECK(hipStreamSynchronize(s));
    shell_l2_errors(m, SbrParams{Omega, a, g, RT}, rho.data(), u.data(), p.data(), out);
    return ATMO_OK;
  });
}

atmo_status atmo_shell_device_bytes(const atmo_shell_desc* shell, const atmo_params* params, size_t* bytes) {
  return guarded(nullptr, [&]() {
    if (!shell || !params || !bytes) throw std::invalid_argument("null argument");
    check_params(*params);
    *bytes = shell_device_bytes(to_shell(shell), to_engine(*params));
    return ATMO_OK;
  });
}

// ------------------------------------------------------------ footprint and ensemble
atmo_status atmo_device_bytes(const atmo_mesh_desc* mesh, const atmo_params* params, size_t* bytes) {
  return guarded(nullptr, [&]() {
    if (!mesh || !params || !bytes) throw std::invalid_argument("null argument");
    check_params(*params);
    *bytes = engine_device_bytes(mesh->degree, to_mesh(mesh), to_engine(*params));
    return ATMO_OK;
  });
}

atmo_status atmo_ctx_device_bytes(atmo_ctx* ctx, size_t* bytes) {
  return guarded(ctx, [&]() {
    if (!ctx || !bytes) throw std::invalid_argument("null argument");
    *bytes = ctx->eng->device_bytes();
    return ATMO_OK;
  });
}

atmo_status atmo_ensemble_create(int slots, double hbm_budget_bytes, int keep_states, atmo_ensemble** out) {
  return guarded(nullptr, [&]() {
    if (!out) throw std::invalid_argument("null argument");
    auto e = std::make_unique<atmo_ensemble>();
    e->ens = std::make_unique<Ensemble>(slots, hbm_budget_bytes, keep_states != 0);
    *out = e.release();
    return ATMO_OK;
  });
}

atmo_status atmo_ensemble_destroy(atmo_ensemble* e) {
  if (!e) return ATMO_INVALID_ARG;
  delete e;
  return ATMO_OK;
}

atmo_status atmo_ensemble_submit(atmo_ensemble* e, const atmo_member_desc* m, int* id) {
  return guarded(nullptr, [&]() {
    if (!e || !m || !id) throw std::invalid_argument("null argument");
    check_params(m->params);
    MemberSpec s;
    s.degree = m->mesh.degree;
    s.mesh = to_mesh(&m->mesh);
    s.params = to_engine(m->params);
    s.theta0 = m->theta0;
    s.p0 = m->p0;
    s.Nbv = m->N;
    s.U = m->U;
    s.g = m->g;
    s.Rgas = m->R;
    s.gamma_gas = m->gamma_gas;
    s.pert = m->perturbation;
    s.seed = m->seed;
    s.nsteps = m->nsteps;
    s.arrival_s = m->arrival_s;
    *id = e->ens->submit(s);
    return ATMO_OK;
  });
}

atmo_status atmo_ensemble_autotune(atmo_ensemble* e, const int* candidates, int ncand, int probe_steps,
                                   int sample_members, double rate_tolerance, int* best_slots, double* rates) {
  return guarded(nullptr, [&]() {
    if (!e || !candidates || ncand < 1 || !best_slots) throw std::invalid_argument("null argument");
    std::vector<double> r;
    const int best = e->ens->autotune(std::vector<int>(candidates, candidates + ncand), probe_steps, sample_members,
                                      rate_tolerance, r);
    *best_slots = best;
    if (rates)
      for (int i = 0; i < ncand; ++i) rates[i] = r[i];
    return ATMO_OK;
  });
}

atmo_status atmo_ensemble_tune_kernels(atmo_ensemble* e, int nrep, atmo_kernel_tuning* out, int cap, int* nclasses) {
  return guarded(nullptr, [&]() {
    if (!e || !nclasses || (cap > 0 && !out)) throw std::invalid_argument("null argument");
    const auto cls = e->ens->tune_kernels(nrep);
    *nclasses = (int)cls.size();
    for (int c = 0; c < (int)cls.size() && c < cap; ++c) to_tuning(cls[c].kt, out + c);
    return ATMO_OK;
  });
}

atmo_status atmo_ensemble_run(atmo_ensemble* e, double* wall_s) {
  return guarded(nullptr, [&]() {
    if (!e) throw std::invalid_argument("null argument");
    e->ens->run();
    if (wall_s) *wall_s = e->ens->wall_s();
    return ATMO_OK;
  });
}

atmo_status atmo_ensemble_result(atmo_ensemble* e, int id, atmo_member_result* out, double* step_ms, int cap) {
  return guarded(nullptr, [&]() {
    if (!e || !out) throw std::invalid_argument("null argument");
    const MemberResult& r = e->ens->result(id);
    out->status = r.status;
    out->arrival_s = r.arrival_s;
    out->admit_s = r.admit_s;
    out->start_s = r.start_s;
    out->finish_s = r.finish_s;
    out->steps_done = (int)r.step_ms.size();
    out->device_bytes = r.device_bytes;
    out->stats = atmo_stats{r.stats.cg_its, r.stats.gmres_its, r.stats.inner_cg_its, r.stats.fp_its,
                            r.stats.mg_cycles, r.stats.coarse_its, r.stats.schur_matvecs, r.stats.host_syncs};
    std::snprintf(out->error, sizeof(out->error), "%s", r.error.c_str());
    if (step_ms)
      for (int i = 0; i < cap && i < (int)r.step_ms.size(); ++i) step_ms[i] = r.step_ms[i];
    return ATMO_OK;
  });
}

atmo_status atmo_ensemble_state(atmo_ensemble* e, int id, double* rho, double* u, double* p, size_t ns) {
  return guarded(nullptr, [&]() {
    if (!e || !rho || !u || !p) throw std::invalid_argument("null argument");
    const MemberResult& r = e->ens->result(id);
    if (r.rho.empty()) throw std::invalid_argument("ensemble: state not kept (keep_states = 0) or member failed");
    if (r.rho.size() != ns) throw std::invalid_argument("ensemble: state size mismatch");
    std::memcpy(rho, r.rho.data(), ns * sizeof(double));
    std::memcpy(u, r.u.data(), 2 * ns * sizeof(double));
    std::memcpy(p, r.p.data(), ns * sizeof(double));
    return ATMO_OK;
  });
}

}  // extern "C"
