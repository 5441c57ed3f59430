// Device engine (see engine.h). Host code here only orchestrates launches on
// one HIP stream; every vector and every Krylov scalar stays in device
// memory. The host reads back a few bytes per Krylov iteration (the
// convergence flag) and the GMRES Hessenberg column.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <type_traits>

#include "amr.h"
#include "dbuf.h"
#include "diag.h"
#include "engine.h"
#include "krylov.h"
#include "mg.h"
#include "schur.h"

namespace atmo {

// Device bytes requested by the DBufs of the calling thread (constructors
// measure their own footprint as the difference around their allocations).
thread_local size_t tl_dev_bytes = 0;

struct Level {
  int nx = 0, nz = 0, ncell = 0;
  long Ns = 0;
  DBuf<double> h;
  MeshDev m{};
  DBuf<double> rho64;
  DBuf<float> rho32;
  DBuf<double> inv64[3];
  DBuf<float> inv32[3];
  double lmax[3] = {1, 1, 1}, lmin[3] = {1, 1, 1};
  DBuf<double> x64, b64, t64, d64;
  DBuf<float> x32, b32, t32, d32;
  DBuf<double> eigb1, eigb2;
  DBuf<double> cg_g, cg_d, cg_h;  // coarse-level CG work
  // locally refined levels: cell geometry, face table (finest), transfer maps
  DBuf<double> geo;
  DBuf<int> faces, childmap, parentmap;
};

// symmetric tridiagonal extreme eigenvalues (Sturm bisection)
static void tridiag_extreme(const std::vector<double>& a, const std::vector<double>& b, double& lmin, double& lmax) {
  const int n = (int)a.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < n; ++i) {
    double r = (i > 0 ? std::abs(b[i - 1]) : 0) + (i + 1 < n ? std::abs(b[i]) : 0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  auto count = [&](double x) {
    int c = 0;
    double q = 1;
    for (int i = 0; i < n; ++i) {
      double bb = i > 0 ? b[i - 1] * b[i - 1] : 0;
      q = a[i] - x - (i > 0 ? bb / q : 0);
      if (q == 0) q = -1e-300;
      if (q < 0) ++c;
    }
    return c;
  };
  auto kth = [&](int k) {
    double l = lo, h = hi;
    for (int it = 0; it < 200; ++it) {
      double mid = 0.5 * (l + h);
      if (count(mid) > k) h = mid;
      else l = mid;
    }
    return 0.5 * (l + h);
  };
  lmin = kth(0);
  lmax = kth(n - 1);
}

__global__ void k_set_cg(CGState* st, double tol, int max_its, double tol_factor) {
  st->tol = tol;
  st->max_its = max_its;
  st->tol_factor = tol_factor;
  st->it = 0;
  st->done = 0;
  st->beta = 0;
}

// st->tol = factor * |b| (deal.II: SolverControl(max_its, eps*rhs.l2_norm()))
__global__ void k_tol_from_norm(CGState* st, const double* nrm, double factor) { st->tol = factor * nrm[0]; }

// extra streams of an engine for the concurrent eigenvalue estimates of its multigrid levels
constexpr int ENGINE_EIG_STREAMS = 3;

template <int K>
class EngineK : public Engine {
 public:
  EngineK(const MeshDesc& mesh, const EngineParams& p) : P(p) {
    const size_t bytes0 = tl_dev_bytes;
    ATMO_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    oc_ = consts();
    build_levels(mesh);
    setup();
    bytes_ = tl_dev_bytes - bytes0;
  }

  // locally refined mesh over the base mesh (amr.h)
  EngineK(const MeshDesc& base, const std::vector<QCell>& leaves, const EngineParams& p) : P(p) {
    const size_t bytes0 = tl_dev_bytes;
    ATMO_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    oc_ = consts();
    build_levels_amr(base, leaves);
    setup();
    bytes_ = tl_dev_bytes - bytes0;
  }

  // buffers, reductions and the geometry-only diagonals of an engine whose levels exist
  void setup() {
    const long Ns = fine().Ns;
    red_partials.alloc(RedBuf::doubles());
    red_partials.zero(s_);
    rb.bind(red_partials.p);
    cgst.alloc(4);
    cgst.zero(s_);
    // eigenvalue estimates: one CG state and coefficient record per level, one reduction buffer per extra stream
    const int nl = (int)lv.size();
    eig_st.alloc(nl);
    eig_st.zero(s_);
    eig_hist.alloc((size_t)2 * MAX_EIG_ITS * nl);
    // (the CG of an estimate reduces one slot at a time: RED_MAX_BLOCKS partials, no ping-pong buffers)
    eig_red.alloc((size_t)EIG_STREAMS * RED_MAX_BLOCKS);
    eig_red.zero(s_);
    eig_concurrent = fine().ncell >= EIG_CONCURRENT_CELLS;
    for (int e = 0; e < EIG_STREAMS; ++e) {
      eig_rb[e].gen = eig_red.p + (size_t)e * RED_MAX_BLOCKS;
      if (!eig_concurrent) continue;
      ATMO_HIP_CHECK(hipStreamCreateWithFlags(&eig_s[e], hipStreamNonBlocking));
      ATMO_HIP_CHECK(hipEventCreateWithFlags(&eig_join[e], hipEventDisableTiming));
    }
    if (eig_concurrent) ATMO_HIP_CHECK(hipEventCreateWithFlags(&eig_fork, hipEventDisableTiming));
    scal.alloc(256);
    scal.zero(s_);
    coarse_its.alloc(1);
    diag_part.alloc(DIAG_PARTIALS);
    diag_rec.alloc(DG_SLOTS);
    diag_rec.zero(s_);
    ATMO_HIP_CHECK(hipHostMalloc((void**)&hst, sizeof(CGState) * 4, hipHostMallocDefault));
    ATMO_HIP_CHECK(hipHostMalloc((void**)&hscal, sizeof(double) * 256, hipHostMallocDefault));
    ATMO_HIP_CHECK(hipHostMalloc((void**)&hint, sizeof(int) * 4, hipHostMallocDefault));
    ATMO_HIP_CHECK(hipHostMalloc((void**)&hgm, sizeof(double) * GM_HOST, hipHostMallocDefault));
    ATMO_HIP_CHECK(hipHostMalloc((void**)&heig, sizeof(double) * 2 * MAX_EIG_ITS * nl, hipHostMallocDefault));
    ATMO_HIP_CHECK(hipHostMalloc((void**)&heig_st, sizeof(CGState) * nl, hipHostMallocDefault));
    for (auto& e : gm_ev) ATMO_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    rho_fix.alloc(Ns);
    pres_fix.alloc(Ns);
    rho_fix.zero(s_);
    pres_fix.zero(s_);
    for (auto* b : {&cg_g, &cg_d, &cg_h, &s_t1, &s_t2, &in_g, &in_d, &in_h}) b->alloc(2 * Ns);
    s_t3.alloc(Ns);
    if (direct_ainv()) ainv.alloc((size_t)fine().ncell * SymBlock<K>::SIZE);
    gm_V.alloc((size_t)(P.gmres_restart + 2) * gm_ld());
    gm_p.alloc(Ns);
    gm_coef.alloc(MAX_GMRES + 2);
    gm_rec.alloc((size_t)MAX_GMRES * GM_REC);
    gm_part.alloc((size_t)RED_MAX_SLOTS * cell_blocks<K>(fine().ncell));
    gm_sig.alloc(MAX_GMRES + 2);
    for (auto* b : {&rho_old, &rho2, &rho3, &rho_curr, &p_old, &p2, &p3, &p_fixed, &p_fixed_old, &rhs_p}) b->alloc(Ns);
    for (auto* b : {&u_old, &u2, &u3, &u_curr, &u_fixed, &tmp1, &rhs_u, &rhs_u0}) b->alloc(2 * Ns);
    for (auto* b : {&rho_old, &rho2, &rho3, &rho_curr, &p_old, &p2, &p3, &p_fixed, &p_fixed_old, &rhs_p, &u_old, &u2,
                    &u3, &u_curr, &u_fixed, &tmp1, &rhs_u, &rhs_u0})
      b->zero(s_);
    // geometry-only diagonals and eigenvalue bounds (M_rho and D blocks)
    for (auto& L : lv) {
      level_diag(*L, BLK_RHO);
      level_diag(*L, BLK_D);
    }
    estimate_eig_levels(BLK_RHO);
    estimate_eig_levels(BLK_D);
    sync();
  }

  ~EngineK() override {
    if (hst) (void)hipHostFree(hst);
    if (hscal) (void)hipHostFree(hscal);
    if (hint) (void)hipHostFree(hint);
    (void)hipStreamSynchronize(s_);
    if (hgm) (void)hipHostFree(hgm);
    if (heig) (void)hipHostFree(heig);
    if (heig_st) (void)hipHostFree(heig_st);
    for (auto& e : gm_ev)
      if (e) (void)hipEventDestroy(e);
    for (int e = 0; e < EIG_STREAMS; ++e) {
      if (eig_s[e]) {
        (void)hipStreamSynchronize(eig_s[e]);
        (void)hipStreamDestroy(eig_s[e]);
      }
      if (eig_join[e]) (void)hipEventDestroy(eig_join[e]);
    }
    if (eig_fork) (void)hipEventDestroy(eig_fork);
    (void)hipStreamDestroy(s_);
  }

  int degree() const override { return K; }
  int ncells() const override { return fine().ncell; }
  long ns() const override { return fine().Ns; }
  int nlevels() const override { return (int)lv.size(); }
  hipStream_t stream() const override { return s_; }
  const EngineParams& params() const override { return P; }
  EngineStats& stats() override { return st_; }

  size_t device_bytes() const override { return bytes_; }
  const std::vector<double>& cell_geometry() const override { return host_cells_; }
  double min_cell_diameter() const override { return amr_min_diam_; }

  void set_params(const EngineParams& p) override {
    const size_t bytes0 = tl_dev_bytes;
    struct Account {
      size_t& total;
      const size_t& now;
      size_t start;
      ~Account() { total += now - start; }
    } account{bytes_, tl_dev_bytes, bytes0};
    const bool prec_change = p.precision != P.precision;
    const bool rebuild = p.gmres_restart != P.gmres_restart;
    P = p;
    oc_ = consts();
    if (rebuild) gm_V.alloc((size_t)(P.gmres_restart + 2) * gm_ld());
    if (prec_change) {
      if (direct_ainv() && !ainv.p) ainv.alloc((size_t)fine().ncell * SymBlock<K>::SIZE);
      // the FP32 copies of the geometry-only diagonals exist only once a form with FP32 levels ran
      for (auto& L : lv) {
        level_diag(*L, BLK_RHO);
        level_diag(*L, BLK_D);
      }
      refresh_levels_rho();
    }
  }

  // ----------------------------------------------------------- operator API
  void set_stages(int hs, int ns) override {
    if (hs < 1 || hs > 3) throw std::invalid_argument("HYPERBOLIC_stage must be 1..3");
    if (ns < 1 || ns > 6) throw std::invalid_argument("NS_stage must be 1..6");
    hyper = hs;
    nsst = ns;
  }

  void set_rho_for_fixed(const double* rho) override {
    vec_copy(fine().Ns, rho, rho_fix.p, s_);
    refresh_levels_rho();
  }

  void set_pres_fixed(const double* p) override { vec_copy(fine().Ns, p, pres_fix.p, s_); }

  void rhs_rho(const double* const* src, int nsrc, double* dst) override {
    if (nsrc != 2 * hyper) throw std::invalid_argument("vmult_rhs_rho_update: expected 2*stage sources");
    launch_rhs_rho<K>(fine().m, oc_, tb_, hyper, src, dst, fine().Ns, s_);
  }
  void rhs_pressure(const double* const* src, int nsrc, double* dst) override {
    const int need = hyper == 1 ? 5 : (hyper == 2 ? 8 : 11);
    if (nsrc != need) throw std::invalid_argument("vmult_rhs_pressure: wrong number of sources");
    launch_rhs_pressure<K>(fine().m, oc_, tb_, hyper, src, dst, fine().Ns, s_);
  }
  void rhs_velocity(const double* const* src, int nsrc, double* dst) override {
    const int need = hyper == 1 ? 4 : (hyper == 2 ? 7 : 9);
    if (nsrc != need) throw std::invalid_argument("vmult_rhs_velocity_fixed: wrong number of sources");
    launch_rhs_velocity<K>(fine().m, oc_, tb_, hyper, src, dst, fine().Ns, s_);
  }
  double bc_coeff() const { return hyper == 1 ? tb_.a22t : tb_.a33t; }
  void vmult_pressure(const double* p, double* dst_u) override {
    launch_B<K>(fine().m, oc_, bc_coeff(), p, dst_u, fine().Ns, s_);
  }
  void vmult_enthalpy(const double* u, double* dst_p) override {
    launch_C<K>(fine().m, oc_, bc_coeff(), pres_fix.p, u, dst_p, fine().Ns, s_);
  }

  void vmult(const double* src, double* dst) override {
    if (nsst == 2) schur_vmult(src, dst);
    else fine_apply(block_of(nsst), src, dst);
  }

  void inverse_diagonal(double* dst) override {
    const int blk = block_of(nsst);
    vec_copy((long)ncomp(blk) * fine().Ns, fine().inv64[blk].p, dst, s_);
  }

  int solve_mass_mg(int ns, double* x, const double* b, double tol_factor) override {
    return pcg_mg(block_of(ns), x, b, tol_factor, P.max_its);
  }

  int solve_schur(double* x, const double* b, double tol_factor) override { return gmres(x, b, tol_factor); }

  // ----------------------------------------------------------- vector operations
  void axpby(long n, double a, const double* x, double b, double* y) override { vec_axpby(n, a, x, b, y, s_); }
  double dot(long n, const double* x, const double* y) override {
    vec_dot(n, x, y, scal.p + 240, false, rb, s_);
    ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 240, scal.p + 240, sizeof(double), hipMemcpyDeviceToHost, s_));
    sync();
    return hscal[240];
  }
  double linf(long n, const double* x) override {
    vec_linf(n, x, nullptr, scal.p + 241, rb, s_);
    ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 241, scal.p + 241, sizeof(double), hipMemcpyDeviceToHost, s_));
    sync();
    return hscal[241];
  }

  // ----------------------------------------------------------- IMEX driver
  void load_state(const double* rho, const double* u, const double* p) override {
    const long Ns = fine().Ns;
    vec_copy(Ns, rho, rho_old.p, s_);
    vec_copy(2 * Ns, u, u_old.p, s_);
    vec_copy(Ns, p, p_old.p, s_);
    tmp1.zero(s_);
  }
  void store_state(double* rho, double* u, double* p) override {
    const long Ns = fine().Ns;
    vec_copy(Ns, rho_old.p, rho, s_);
    vec_copy(2 * Ns, u_old.p, u, s_);
    vec_copy(Ns, p_old.p, p, s_);
  }
  double* state_ptr(int which) override { return which == 0 ? rho_old.p : (which == 1 ? u_old.p : p_old.p); }

  // The tunable kernel shapes timed on this mesh (two alternating rounds, the faster round of each
  // candidate counts): the Arnoldi kernel's waves per SIMD (structured meshes of the cellwise-direct forms)
  // and the FP32 sweeps' form (matrix cores, 16 cells per wave, or vector ALUs, 64 / (2k+1) cells per wave).
  KernelTuning tune_kernels(int nrep) override {
    KernelTuning kt;
    double ms = 0, bytes = 0;
    if (direct_ainv() && !fine().m.cface && P.gmres_restart >= 13) {
      kt.arnoldi_us[0] = kt.arnoldi_us[1] = 1e300;
      for (int round = 0; round < 2; ++round)
        for (int i = 0; i < 2; ++i) {
          P.arnoldi_waves = 3 + i;
          time_kernel(12, nrep, &ms, &bytes);
          kt.arnoldi_us[i] = std::min(kt.arnoldi_us[i], 1e3 * ms);
        }
      kt.arnoldi_waves = kt.arnoldi_us[0] < kt.arnoldi_us[1] ? 3 : 4;
      P.arnoldi_waves = kt.arnoldi_waves;
    }
    if (mg32() && (K == 2 || K == 3)) {
      kt.sweep_us[0] = kt.sweep_us[1] = 1e300;
      for (int round = 0; round < 2; ++round)
        for (int i = 0; i < 2; ++i) {
          time_kernel(i ? 14 : 15, nrep, &ms, &bytes);
          kt.sweep_us[i] = std::min(kt.sweep_us[i], 1e3 * ms);
        }
      kt.mg_matrix_cores = kt.sweep_us[1] <= kt.sweep_us[0] ? 1 : 0;
      P.mg_matrix_cores = kt.mg_matrix_cores;
    }
    return kt;
  }

  void time_kernel(int which, int nrep, double* ms, double* bytes) override {
    const long Ns = fine().Ns;
    const MeshDev& m = fine().m;
    double fields = 0;
    hipEvent_t e0, e1;
    ATMO_HIP_CHECK(hipEventCreate(&e0));
    ATMO_HIP_CHECK(hipEventCreate(&e1));
    CGState* st = cgst.p + 3;
    if (which == 1 || which == 6) {
      hipLaunchKernelGGL(k_set_cg, dim3(1), dim3(1), 0, s_, st, 0.0, 1 << 30, 1e-300);
      vec_copy(2 * Ns, u_old.p, s_t1.p, s_);
      vec_fill(2 * Ns, 0.0, s_t2.p, s_);
      jcg_init(Ns, m, K, s_t1.p, fine().inv64[BLK_A].p, in_g.p, in_d.p, st, rb, s_);
      jcg_apply<K>(1, m, rho_fix.p, fine().inv64[BLK_A].p, in_g.p, in_d.p, in_h.p, st, rb, Ns, s_);
    }
    ChebSteps cs_t;  // sweep coefficients for the timing of 14/15 (the kernels' work does not depend on them)
    if (which == 14 || which == 15) {
      if (!mg32()) throw std::invalid_argument("time_kernel 14/15 need the mixed form (precision 1)");
      cs_t.n = P.cheb_degree;
      for (int s = 0; s < cs_t.n; ++s) {
        cs_t.f1[s] = s ? 0.5 : 0.0;
        cs_t.f2[s] = 0.25;
      }
      vec_cast<float, double>(2 * Ns, u_old.p, fine().b32.p, s_);
    }
    auto launch = [&]() {
      switch (which) {
        case 14:  // FP32 multigrid pre-smoothing sweep + residual of the A block on the finest level: matrix cores
        case 15:  // the same sweep on the vector ALUs
        {
          Level& F = fine();
          launch_presmooth_residual<K, float>(F.m, 2, true, F.x32.p, F.rho32.p, (float)coeff(BLK_A), F.b32.p,
                                              F.inv32[BLK_A].p, cs_t, F.t32.p, Ns, which == 14, s_);
          fields = 0.5 * 9;  // FP32: rho, b (2), P (2) in; x (2), res (2) out
          break;
        }
        case 0:
          launch_mass<K, double>(m, 2, true, u_old.p, rho_fix.p, in_h.p, 1.0, nullptr, 0.0, Ns, s_);
          fields = 5;
          break;
        case 1:
          jcg_apply<K>(1, m, rho_fix.p, fine().inv64[BLK_A].p, in_g.p, in_d.p, in_h.p, st, rb, Ns, s_);
          fields = 11;
          break;
        case 6:
          jcg_update(1, Ns, m, K, s_t2.p, in_g.p, in_d.p, in_h.p, fine().inv64[BLK_A].p, st, rb, s_);
          fields = 14;
          break;
        case 7:  // plain device copy of 2*Ns doubles (counter calibration)
          ATMO_HIP_CHECK(hipMemcpyAsync(s_t1.p, u_old.p, 2 * Ns * sizeof(double), hipMemcpyDeviceToDevice, s_));
          fields = 4;
          break;
        case 2:
          launch_mass<K, double>(m, 1, false, rho_old.p, nullptr, gm_p.p, 1.0, nullptr, 0.0, Ns, s_);
          fields = 2;
          break;
        case 3:
          launch_B<K>(m, oc_, bc_coeff(), p_old.p, in_h.p, Ns, s_);
          fields = 3;
          break;
        case 4:
          launch_C<K>(m, oc_, bc_coeff(), pres_fix.p, u_old.p, gm_p.p, Ns, s_);
          fields = 4;
          break;
        case 5: {
          const double* sp[11] = {rho_old.p, u_old.p, p_old.p, rho_old.p, u_old.p, p_old.p,
                                  rho_old.p, u_old.p, p_old.p, rho_old.p, u_old.p};
          launch_rhs_pressure<K>(m, oc_, tb_, 3, sp, gm_p.p, Ns, s_);
          fields = 16;
          break;
        }
        case 8:  // mixed Schur: B and the cellwise A^{-1} in one pass
          if (!ainv.p) throw std::invalid_argument("time_kernel 8 needs a cellwise-direct A^{-1} (precision 1 or 2)");
          launch_b_ainv<K>(m, oc_, bc_coeff(), p_old.p, ainv.p, in_h.p, Ns, s_);
          // p, y (2 components), and the packed inverse block per cell
          fields = 3 + (double)SymBlock<K>::SIZE / ((K + 1) * (K + 1));
          break;
        case 9:  // D p - C u, scaled (GMRES's preconditioned Schur product)
          launch_d_minus_c<K>(m, oc_, bc_coeff(), pres_fix.p, u_old.p, p_old.p, fine().inv64[BLK_D].p, gm_p.p, Ns,
                              s_);
          fields = 6;
          break;
        case 10:  // GMRES dot products: w = V_13 against V_0 .. V_13 (14 columns, w included)
        case 11:  // GMRES projection + norm: V_13 -= sum of 13 columns
          if (P.gmres_restart < 13) throw std::invalid_argument("time_kernel 10/11 need gmres_restart >= 13");
          if (which == 10) {
            gm_mdot(Ns, gm_V.p + 13 * gm_ld(), gm_V.p, gm_ld(), 14, gm_rec.p, rb, s_);
            fields = 15;
          } else {
            gm_maxpy_norm(Ns, gm_V.p + 13 * gm_ld(), gm_V.p, gm_ld(), 13, gm_rec.p, gm_sig.p, gm_rec.p + GM_REC - 1,
                          gm_sig.p + 14, rb, s_);
            fields = 15;
          }
          break;
        case 12:  // the Arnoldi step's scaled D - C with the dot products against 14 basis vectors
        case 13:  // the same kernel with no basis vector (D - C and w.w only)
        {
          if (P.gmres_restart < 13) throw std::invalid_argument("time_kernel 12/13 need gmres_restart >= 13");
          // as GMRES runs it: the operator applied to the newest basis vector (V_13), whose staged values also
          // serve its own dot product (p, y, p_fix, the Jacobi scale, 13 other basis vectors in; w out)
          const int nv = which == 12 ? 14 : 0;
          const int G = cell_blocks<K>(fine().ncell);
          const double* p = which == 12 ? gm_V.p + 13 * gm_ld() : p_old.p;
          launch_d_minus_c_dots<K>(m, oc_, bc_coeff(), pres_fix.p, u_old.p, p, fine().inv64[BLK_D].p,
                                   gm_V.p + 14 * gm_ld(), Ns, gm_V.p, gm_ld(), nv, gm_part.p, G, P.arnoldi_waves, s_);
          fields = nv ? 6 + nv - 1 : 6;
          break;
        }
        default:
          throw std::invalid_argument("time_kernel: unknown kernel id");
      }
    };
    launch();  // warm-up
    ATMO_HIP_CHECK(hipEventRecord(e0, s_));
    for (int r = 0; r < nrep; ++r) launch();
    ATMO_HIP_CHECK(hipEventRecord(e1, s_));
    ATMO_HIP_CHECK(hipEventSynchronize(e1));
    float total = 0;
    ATMO_HIP_CHECK(hipEventElapsedTime(&total, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)total / nrep;
    *bytes = fields * 8.0 * (double)Ns;
  }

  // one IMEX step (Atmospheric_Flow.cc:1115-1222)
  void step() override {
    const long Ns = fine().Ns;
    // stage 1
    update_density(1);
    launch_diag_density<K>(fine().m, rho2.p, 0, diag_part.p, diag_rec.p, s_);
    set_rho_for_fixed(rho2.p);
    vec_copy(Ns, p_old.p, p_fixed_old.p, s_);
    vec_copy(2 * Ns, u_old.p, u_fixed.p, s_);
    fixed_point_loop(1);
    vec_copy(Ns, p_fixed.p, p2.p, s_);
    vec_copy(2 * Ns, u_fixed.p, u2.p, s_);
    // stage 2
    update_density(2);
    launch_diag_density<K>(fine().m, rho3.p, 1, diag_part.p, diag_rec.p, s_);
    set_rho_for_fixed(rho3.p);
    vec_copy(Ns, p2.p, p_fixed_old.p, s_);
    vec_copy(2 * Ns, u2.p, u_fixed.p, s_);
    fixed_point_loop(2);
    vec_copy(Ns, p_fixed.p, p3.p, s_);
    vec_copy(2 * Ns, u_fixed.p, u3.p, s_);
    // stage 3
    update_density(3);
    launch_diag_density<K>(fine().m, rho_curr.p, 2, diag_part.p, diag_rec.p, s_);
    set_rho_for_fixed(rho_curr.p);
    update_velocity(3);
    update_pressure();
    vec_copy(Ns, rho_curr.p, rho_old.p, s_);
    vec_copy(2 * Ns, u_curr.p, u_old.p, s_);
    launch_diag_flow<K>(fine().m, rho_old.p, u_old.p, p_old.p, P.gamma_gas, Ns, diag_part.p, diag_rec.p, s_);
  }

  void step_diagnostics(double* rec, bool of_state) override {
    if (of_state) {
      launch_diag_density<K>(fine().m, rho_old.p, 2, diag_part.p, diag_rec.p, s_);
      launch_diag_flow<K>(fine().m, rho_old.p, u_old.p, p_old.p, P.gamma_gas, fine().Ns, diag_part.p, diag_rec.p,
                          s_);
    }
    ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 128, diag_rec.p, sizeof(double) * DG_SLOTS, hipMemcpyDeviceToHost, s_));
    sync();
    std::memcpy(rec, hscal + 128, sizeof(double) * DG_SLOTS);
  }

 private:
  EngineParams P;
  Tableau tb_;
  OpConsts oc_{};
  hipStream_t s_ = nullptr;
  std::vector<std::unique_ptr<Level>> lv;
  RedBuf rb;
  DBuf<double> red_partials;
  DBuf<double> diag_part, diag_rec;  // per-step diagnostics: block partials and the step record
  DBuf<double> ainv;  // FP64 inverse cell blocks of A, packed symmetric (mixed form)
  DBuf<CGState> cgst;
  DBuf<double> scal;
  DBuf<int> coarse_its;
  CGState* hst = nullptr;
  double* hscal = nullptr;
  int* hint = nullptr;
  // GMRES: pinned host copies of the per-iteration records followed by the
  // update coefficients, and one event per iteration
  static constexpr int GM_HOST = MAX_GMRES * GM_REC + MAX_GMRES + 2;
  double* hgm = nullptr;
  // Chebyshev eigenvalue estimates of all levels run concurrently: the finest level on s_, the others
  // spread over EIG_STREAMS extra streams (default priority), each with its own reduction buffer
  static constexpr int EIG_STREAMS = ENGINE_EIG_STREAMS;
  // below this many fine cells the estimates are launch-bound and stay on s_: with many engines on one
  // GPU (ensembles) the extra streams of each would share the few hardware queues
  static constexpr int EIG_CONCURRENT_CELLS = 65536;
  bool eig_concurrent = false;
  hipStream_t eig_s[EIG_STREAMS] = {};
  hipEvent_t eig_fork = nullptr, eig_join[EIG_STREAMS] = {};
  RedBuf eig_rb[EIG_STREAMS];
  DBuf<double> eig_red;
  DBuf<CGState> eig_st;      // per level
  CGState* heig_st = nullptr;  // pinned copy of eig_st
  double* heig = nullptr;      // pinned copy of eig_hist
  DBuf<double> eig_hist;       // per level: CG coefficients of the eigenvalue estimate, alpha_it, beta_it
  hipEvent_t gm_ev[MAX_GMRES] = {};
  DBuf<double> gm_rec, gm_sig;
  DBuf<double> gm_part;  // per-block partial dot products of the fused Arnoldi step, [slot][block]
  int hyper = 1, nsst = 1;
  DBuf<double> rho_fix, pres_fix;
  DBuf<double> cg_g, cg_d, cg_h, s_t1, s_t2, s_t3, in_g, in_d, in_h, gm_V, gm_p, gm_coef;
  DBuf<double> rho_old, rho2, rho3, rho_curr, u_old, u2, u3, u_curr, u_fixed, p_old, p2, p3, p_fixed, p_fixed_old,
      tmp1, rhs_u, rhs_p;
  DBuf<double> rhs_u0;  // rhs_u of the current stage before update_velocity subtracts B p_fixed
  EngineStats st_;
  int last_inner = 8;  // iterations of the last inner CG (first batch of the next)
  size_t bytes_ = 0;  // device bytes requested by this engine's buffers
  double amr_min_diam_ = 0.0;      // locally refined mesh: minimal cell diameter
  std::vector<double> host_cells_;  // locally refined mesh: finest-level cell geometry (6 per cell)

  OpConsts consts() const {
    OpConsts o;
    o.Ma = P.Ma;
    o.Fr = P.Fr;
    o.dt = P.dt;
    o.gamma_gas = P.gamma_gas;
    o.grav[0] = P.grav[0];
    o.grav[1] = P.grav[1];
    o.grav_factor = P.grav_factor;
    return o;
  }

  Level& fine() { return *lv.back(); }
  const Level& fine() const { return *lv.back(); }
  // precision 1 (mixed) and 2 (its FP64 comparator) invert A cell by cell; only 1 runs the multigrid in FP32
  bool direct_ainv() const { return P.precision == 1 || P.precision == 2; }
  bool mg32() const { return P.precision == 1; }
  static int block_of(int ns) { return (ns == 1 || ns == 4) ? BLK_RHO : ((ns == 3 || ns == 5) ? BLK_A : BLK_D); }
  static int ncomp(int blk) { return blk == BLK_A ? 2 : 1; }
  double coeff(int blk) const { return blk == BLK_D ? 1.0 / (P.gamma_gas - 1.0) : 1.0; }

  void sync() {
    ATMO_HIP_CHECK(hipStreamSynchronize(s_));
    ++st_.host_syncs;
  }

  void build_levels(const MeshDesc& mesh) {
    if (mesh.nx < 1 || mesh.nz < 1 || (int)mesh.h.size() != mesh.nx + 1)
      throw std::invalid_argument("mesh: need nx, nz >= 1 and nx+1 terrain heights");
    struct LD {
      int nx, nz;
      std::vector<double> h;
    };
    std::vector<LD> ls;
    ls.push_back({mesh.nx, mesh.nz, mesh.h});
    while (ls.back().nx % 2 == 0 && ls.back().nz % 2 == 0 && ls.back().nx >= 4 &&
           (ls.back().nx * ls.back().nz) / 4 >= P.mg_min_cells) {
      LD c;
      c.nx = ls.back().nx / 2;
      c.nz = ls.back().nz / 2;
      c.h.resize(c.nx + 1);
      for (int i = 0; i <= c.nx; ++i) c.h[i] = ls.back().h[2 * i];
      ls.push_back(c);
    }
    const int nlev = (int)ls.size();
    for (int l = 0; l < nlev; ++l) {
      const LD& d = ls[nlev - 1 - l];
      auto L = std::make_unique<Level>();
      L->nx = d.nx;
      L->nz = d.nz;
      L->ncell = d.nx * d.nz;
      L->Ns = (long)L->ncell * (K + 1) * (K + 1);
      L->h.alloc(d.nx + 1);
      ATMO_HIP_CHECK(hipMemcpy(L->h.p, d.h.data(), sizeof(double) * (d.nx + 1), hipMemcpyHostToDevice));
      L->m.h = L->h.p;
      L->m.nx = d.nx;
      L->m.nz = d.nz;
      L->m.dx = mesh.Lx / d.nx;
      L->m.H = mesh.H;
      L->m.inv_nz = 1.0 / d.nz;
      L->m.periodic = mesh.periodic;
      L->m.ncell = L->ncell;
      alloc_level_vectors(*L, l == 0);
      lv.push_back(std::move(L));
    }
  }

  // work vectors of one level (coarse: with the coarse-CG buffers)
  void alloc_level_vectors(Level& L, bool coarse) {
    const long n2 = 2 * L.Ns;
    L.rho64.alloc(L.Ns);
    L.rho32.alloc(L.Ns);
    L.rho64.zero(s_);
    L.rho32.zero(s_);
    for (int b = 0; b < 3; ++b) {
      L.inv64[b].alloc(b == BLK_A ? n2 : L.Ns);
      L.inv32[b].alloc(b == BLK_A ? n2 : L.Ns);
    }
    for (auto* b : {&L.x64, &L.b64, &L.t64, &L.d64}) b->alloc(n2);
    for (auto* b : {&L.x32, &L.b32, &L.t32, &L.d32}) b->alloc(n2);
    if (coarse)
      for (auto* b : {&L.cg_g, &L.cg_d, &L.cg_h}) b->alloc(n2);
    // Chebyshev eigenvalue start vectors: (i % 11) - mean
    for (int nc = 1; nc <= 2; ++nc) {
      const long n = nc * L.Ns;
      std::vector<double> v(n);
      double mean = 0;
      for (long i = 0; i < n; ++i) {
        v[i] = (double)(i % 11);
        mean += v[i];
      }
      mean /= (double)n;
      for (auto& x : v) x -= mean;
      DBuf<double>& dst = nc == 1 ? L.eigb1 : L.eigb2;
      dst.alloc(n);
      ATMO_HIP_CHECK(hipMemcpy(dst.p, v.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
  }

  // Levels of a locally refined mesh: the uniformly coarsened base mesh
  // below, the globally coarsened quadtree levels above (amr.h).
  void build_levels_amr(const MeshDesc& base, const std::vector<QCell>& leaves) {
    std::vector<AmrLevelHost> hl = build_amr_levels(base, leaves);
    amr_min_diam_ = amr_min_cell_diameter(hl.front());
    build_levels(base);  // structured levels, base mesh finest
    if (hl.size() == 1) return;  // nothing refined: the base mesh is the mesh
    auto upload_int = [&](DBuf<int>& d, const std::vector<int>& v) {
      d.alloc(v.size());
      ATMO_HIP_CHECK(hipMemcpy(d.p, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice));
    };
    // the base level's children live in the coarsest quadtree level above it
    upload_int(lv.back()->childmap, hl[hl.size() - 1].child);
    lv.back()->m.child = lv.back()->childmap.p;
    for (int i = (int)hl.size() - 2; i >= 0; --i) {
      const AmrLevelHost& H = hl[i];
      auto L = std::make_unique<Level>();
      L->ncell = H.ncell;
      L->nx = H.ncell;
      L->nz = 1;
      L->Ns = (long)H.ncell * (K + 1) * (K + 1);
      L->geo.alloc(H.cgeo.size());
      ATMO_HIP_CHECK(hipMemcpy(L->geo.p, H.cgeo.data(), sizeof(double) * H.cgeo.size(), hipMemcpyHostToDevice));
      upload_int(L->parentmap, H.parent);
      if (i > 0) upload_int(L->childmap, H.child);
      if (i == 0) upload_int(L->faces, H.cface);
      L->m.nx = H.ncell;
      L->m.nz = 1;
      L->m.dx = 0.0;
      L->m.H = base.H;
      L->m.inv_nz = 1.0;
      L->m.periodic = base.periodic;
      L->m.ncell = H.ncell;
      L->m.cgeo = L->geo.p;
      L->m.parent = L->parentmap.p;
      L->m.child = i > 0 ? L->childmap.p : nullptr;
      L->m.cface = i == 0 ? L->faces.p : nullptr;
      alloc_level_vectors(*L, false);
      lv.push_back(std::move(L));
    }
    host_cells_ = hl.front().cgeo;
  }

  // level operator in precision T: dst = Op_blk src (+ add_scale*add)
  template <typename T>
  void level_apply(Level& L, int blk, const T* src, T* dst, T cf = 1, const T* add = nullptr, T add_scale = 0,
                   hipStream_t s = nullptr);

  void fine_apply(int blk, const double* src, double* dst) { level_apply<double>(fine(), blk, src, dst); }

  void level_diag(Level& L, int blk) {
    const bool w = blk == BLK_A;
    launch_inv_diag<K, double>(L.m, ncomp(blk), w, L.rho64.p, L.inv64[blk].p, coeff(blk), L.Ns, s_);
    if (mg32()) vec_cast<float, double>((long)ncomp(blk) * L.Ns, L.inv64[blk].p, L.inv32[blk].p, s_);
  }

  void refresh_levels_rho() {
    Level& F = fine();
    vec_copy(F.Ns, rho_fix.p, F.rho64.p, s_);
    for (int l = (int)lv.size() - 1; l > 0; --l) launch_project_down<K>(lv[l]->m, lv[l - 1]->m, lv[l]->rho64.p, lv[l - 1]->rho64.p, s_);
    for (auto& L : lv) {
      if (mg32()) vec_cast<float, double>(L->Ns, L->rho64.p, L->rho32.p, s_);
      level_diag(*L, BLK_A);
    }
    estimate_eig_levels(BLK_A);
    if (direct_ainv()) build_block_inverse<K>(F.m, rho_fix.p, ainv.p, s_);
  }

  // Chebyshev eigenvalue estimate: <= cheb_eig_its CG iterations with the
  // Jacobi preconditioner from b = (i % 11) - mean, Lanczos tridiagonal,
  // lmax = 1.2 * estimate, lmin = lmax / range.
  //
  // Every iteration is queued without a host round trip: the CG kernels
  // are no-ops once the device-side state says the solve has finished, and
  // each iteration's alpha and beta are recorded on the device (eig_hist);
  // the host reads them back once. The iterations that ran are the first
  // st.it (the one that finished the solve included).
  void eig_enqueue(int l, int blk, hipStream_t s, RedBuf& r) {
    Level& L = *lv[l];
    const long n = (long)ncomp(blk) * L.Ns;
    const double* b = ncomp(blk) == 1 ? L.eigb1.p : L.eigb2.p;
    CGState* st = eig_st.p + l;
    double *g = L.x64.p, *d = L.b64.p, *h = L.t64.p, *x = L.d64.p;
    const double* Pd = L.inv64[blk].p;
    double* hist = eig_hist.p + (size_t)2 * MAX_EIG_ITS * l;
    const int nit = P.cheb_eig_its;
    hipLaunchKernelGGL(k_set_cg, dim3(1), dim3(1), 0, s, st, 1e-2, nit, 0.0);
    cg_residual(n, g, b, false, s);
    vec_fill(n, 0.0, x, s);
    vec_mul(n, g, Pd, h, s);
    vec_axpby(n, -1.0, h, 0.0, d, s);
    cg_dot(n, g, h, st, DOT_GH_INIT, r, s);
    for (int it = 0; it < nit; ++it) {
      level_apply<double>(L, blk, d, h, 1.0, nullptr, 0.0, s);
      cg_dot(n, d, h, st, DOT_ALPHA, r, s, hist + 2 * it);
      cg_update(n, x, g, d, h, st, r, s);
      if (it + 1 == nit) break;  // the last beta is never used
      vec_mul(n, g, Pd, h, s);
      cg_dot(n, g, h, st, DOT_BETA, r, s, hist + 2 * it + 1);
      cg_dir(n, d, h, st, s);
    }
  }

  void eig_bounds(int l, int blk) {
    Level& L = *lv[l];
    const int nit = P.cheb_eig_its;
    const double* he = heig + (size_t)2 * MAX_EIG_ITS * l;
    const int m = std::min(nit, heig_st[l].it);
    std::vector<double> a(m), o(m > 0 ? m - 1 : 0);
    for (int j = 0; j < m; ++j) {
      a[j] = 1.0 / he[2 * j] + (j > 0 ? he[2 * (j - 1) + 1] / he[2 * (j - 1)] : 0.0);
      if (j + 1 < m) o[j] = std::sqrt(he[2 * j + 1]) / he[2 * j];
    }
    double lmin, lmax;
    tridiag_extreme(a, o, lmin, lmax);
    L.lmax[blk] = 1.2 * lmax;
    L.lmin[blk] = L.lmax[blk] / P.cheb_range;
  }

  // The estimates of levels 1 .. finest, each on its own level's vectors. Levels of at most
  // EIG_SMALL_CELLS cells run together in one launch (a workgroup each, all iterations inside: about 90
  // launches of tiny kernels per level otherwise); the others kernel by kernel, the finest on s_. With
  // concurrent estimates the batch and the other levels go to the extra streams, which start after the
  // work queued on s_ so far and which s_ waits for.
  static constexpr int EIG_SMALL_CELLS = 512;
  void eig_enqueue_levels(int blk) {
    const int nl = (int)lv.size();
    EigBatch B{};
    B.n = 0;
    std::vector<int> big;
    for (int l = 1; l < nl; ++l) {
      Level& L = *lv[l];
      if (L.ncell <= EIG_SMALL_CELLS && B.n < EIG_BATCH_MAX) {
        EigLevel& e = B.L[B.n++];
        e.m = L.m;
        e.rho = blk == BLK_A ? L.rho64.p : nullptr;
        e.P = L.inv64[blk].p;
        e.b = ncomp(blk) == 1 ? L.eigb1.p : L.eigb2.p;
        e.g = L.x64.p;
        e.d = L.b64.p;
        e.h = L.t64.p;
        e.x = L.d64.p;
        e.hist = eig_hist.p + (size_t)2 * MAX_EIG_ITS * l;
        e.st = eig_st.p + l;
        e.Ns = L.Ns;
      } else {
        big.push_back(l);
      }
    }
    const double eig_tol = 1e-2;  // absolute, on |g|
    if (!eig_concurrent) {
      launch_eig_cg_batch<K>(B, ncomp(blk), blk == BLK_A, coeff(blk), P.cheb_eig_its, eig_tol, s_);
      for (int l : big) eig_enqueue(l, blk, s_, rb);
      return;
    }
    ATMO_HIP_CHECK(hipEventRecord(eig_fork, s_));
    for (int e = 0; e < EIG_STREAMS; ++e) ATMO_HIP_CHECK(hipStreamWaitEvent(eig_s[e], eig_fork, 0));
    launch_eig_cg_batch<K>(B, ncomp(blk), blk == BLK_A, coeff(blk), P.cheb_eig_its, eig_tol, eig_s[0]);
    int e = 1;
    for (int l : big) {
      if (l == nl - 1) {
        eig_enqueue(l, blk, s_, rb);
      } else {
        eig_enqueue(l, blk, eig_s[e], eig_rb[e]);
        e = e % (EIG_STREAMS - 1) + 1;
      }
    }
    for (int e2 = 0; e2 < EIG_STREAMS; ++e2) {
      ATMO_HIP_CHECK(hipEventRecord(eig_join[e2], eig_s[e2]));
      ATMO_HIP_CHECK(hipStreamWaitEvent(s_, eig_join[e2], 0));
    }
  }

  // Bounds of every level for block blk; one host synchronization reads every level's record. (Capturing
  // the A block's estimates into a HIP graph saved 1.4 % of the mixed step but failed when several engines
  // run in separate threads: another thread's allocation or free during a capture invalidates it.)
  void estimate_eig_levels(int blk) {
    const int nl = (int)lv.size();
    if (nl < 2) return;
    eig_enqueue_levels(blk);
    ATMO_HIP_CHECK(hipMemcpyAsync(heig, eig_hist.p, sizeof(double) * 2 * MAX_EIG_ITS * nl, hipMemcpyDeviceToHost, s_));
    ATMO_HIP_CHECK(hipMemcpyAsync(heig_st, eig_st.p, sizeof(CGState) * nl, hipMemcpyDeviceToHost, s_));
    sync();
    for (int l = 1; l < nl; ++l) eig_bounds(l, blk);
  }

  void read_state(int slot) {
    ATMO_HIP_CHECK(hipMemcpyAsync(&hst[slot], cgst.p + slot, sizeof(CGState), hipMemcpyDeviceToHost, s_));
    sync();
  }

  // --------------------------------------------------------------- MG
  template <typename T>
  static T* LX(Level& L) {
    if constexpr (std::is_same<T, double>::value) return L.x64.p; else return L.x32.p;
  }
  template <typename T>
  static T* LB(Level& L) {
    if constexpr (std::is_same<T, double>::value) return L.b64.p; else return L.b32.p;
  }
  template <typename T>
  static T* LT(Level& L) {
    if constexpr (std::is_same<T, double>::value) return L.t64.p; else return L.t32.p;
  }
  template <typename T>
  static T* LD(Level& L) {
    if constexpr (std::is_same<T, double>::value) return L.d64.p; else return L.d32.p;
  }
  template <typename T>
  static const T* LP(Level& L, int blk) {
    if constexpr (std::is_same<T, double>::value) return L.inv64[blk].p; else return L.inv32[blk].p;
  }

  // Coefficients of one PreconditionChebyshev sweep of degree cheb_degree on
  // [lmin, lmax] of level l: step 0 is d = P (b - A x) / theta, step k the
  // three-term recurrence.
  ChebSteps cheb_steps(int l, int blk) const {
    const Level& L = *lv[l];
    const double theta = 0.5 * (L.lmax[blk] + L.lmin[blk]), delta = 0.5 * (L.lmax[blk] - L.lmin[blk]);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    ChebSteps cs;
    cs.n = P.cheb_degree;
    cs.f1[0] = 0.0;
    cs.f2[0] = 1.0 / theta;
    for (int k = 1; k < cs.n; ++k) {
      const double rn = 1.0 / (2.0 * sigma - rho);
      cs.f1[k] = rn * rho;
      cs.f2[k] = 2.0 * rn / delta;
      rho = rn;
    }
    return cs;
  }

  // V-cycle on level l: x = MG(b); x, b are level buffers of type T
  template <typename T>
  void vcycle(int l, int blk, T* x, const T* b, const double* tol, int max_its) {
    Level& L = *lv[l];
    const int nc = ncomp(blk);
    if (l == 0) {
      if constexpr (std::is_same<T, double>::value)
        launch_coarse_cg<K>(L.m, nc, blk == BLK_A, L.rho64.p, coeff(blk), b, x, L.cg_g.p, L.cg_d.p, L.cg_h.p, tol,
                            max_its, coarse_its.p, L.Ns, s_);
      else
        launch_coarse_cg_f32<K>(L.m, nc, blk == BLK_A, L.rho64.p, coeff(blk), b, x, L.b64.p, L.x64.p, L.cg_g.p,
                                L.cg_d.p, L.cg_h.p, tol, max_its, coarse_its.p, L.Ns, s_);
      return;
    }
    // pre-smoothing from a zero guess and the residual t = b - A x: one kernel
    T* t = LT<T>(L);
    const T* Pd = LP<T>(L, blk);
    const T* rw = nullptr;
    if (blk == BLK_A) rw = std::is_same<T, double>::value ? (const T*)L.rho64.p : (const T*)L.rho32.p;
    const ChebSteps cs = cheb_steps(l, blk);
    launch_presmooth_residual<K, T>(L.m, nc, blk == BLK_A, x, rw, (T)coeff(blk), b, Pd, cs, t, L.Ns, P.mg_matrix_cores,
                                    s_);
    Level& C = *lv[l - 1];
    launch_restrict<K, T>(L.m, C.m, t, LB<T>(C), nc, L.Ns, C.Ns, s_);
    vcycle<T>(l - 1, blk, LX<T>(C), LB<T>(C), tol, max_its);
    // coarse-grid correction and post-smoothing: one kernel
    launch_prolongate_postsmooth<K, T>(L.m, C.m, nc, blk == BLK_A, x, LX<T>(C), rw, (T)coeff(blk), b, Pd, cs, L.Ns,
                                       C.Ns, P.mg_matrix_cores, s_);
  }

  // PreconditionMG::vmult: dst = V-cycle(src) on the fine level
  void mg_prec(int blk, const double* src, double* dst, const double* tol, int max_its) {
    ++st_.mg_cycles;
    const int top = (int)lv.size() - 1;
    if (mg32()) {
      Level& F = fine();
      const long n = (long)ncomp(blk) * F.Ns;
      vec_cast<float, double>(n, src, F.b32.p, s_);
      vcycle<float>(top, blk, F.x32.p, F.b32.p, tol, max_its);
      vec_cast<double, float>(n, F.x32.p, dst, s_);
    } else {
      vcycle<double>(top, blk, dst, src, tol, max_its);
    }
  }

  // CG + MG on a mass-type block (SolverCG + PreconditionMG). Tolerance
  // tol_factor*|b|; the coarse CG shares the tolerance and max_its
  // (Atmospheric_Flow.cc:452/506 share one SolverControl).
  int pcg_mg(int blk, double* x, const double* b, double tol_factor, int max_its) {
    const long n = (long)ncomp(blk) * fine().Ns;
    CGState* st = cgst.p + 0;
    hipLaunchKernelGGL(k_set_cg, dim3(1), dim3(1), 0, s_, st, 0.0, max_its, tol_factor);
    vec_dot(n, b, b, scal.p + 0, true, rb, s_);
    hipLaunchKernelGGL(k_tol_from_norm, dim3(1), dim3(1), 0, s_, st, scal.p + 0, tol_factor);
    double *g = cg_g.p, *d = cg_d.p, *h = cg_h.p;
    fine_apply(blk, x, g);
    cg_residual(n, g, b, true, s_);
    cg_init_check(n, g, st, false, rb, s_);
    read_state(0);
    if (hst[0].done == 1) return 0;
    if (hst[0].done == -1) throw NoConvergence("CG+MG: no convergence");
    const double* tol = &st->tol;
    mg_prec(blk, g, h, tol, max_its);
    vec_axpby(n, -1.0, h, 0.0, d, s_);
    cg_dot(n, g, h, st, DOT_GH_INIT, rb, s_);
    int it = 0;
    while (true) {
      fine_apply(blk, d, h);
      cg_dot(n, d, h, st, DOT_ALPHA, rb, s_);
      cg_update(n, x, g, d, h, st, rb, s_);
      ATMO_HIP_CHECK(hipMemcpyAsync(hint, coarse_its.p, sizeof(int), hipMemcpyDeviceToHost, s_));
      read_state(0);
      if (hint[0] < 0) throw NoConvergence("MG coarse CG: no convergence");
      st_.coarse_its += hint[0];
      it = hst[0].it;
      if (hst[0].done == 1) break;
      if (hst[0].done == -1) throw NoConvergence("CG+MG: no convergence");
      mg_prec(blk, g, h, tol, max_its);
      cg_dot(n, g, h, st, DOT_BETA, rb, s_);
      cg_dir(n, d, h, st, s_);
    }
    st_.cg_its += it;
    return it;
  }

  // inner A^{-1} of the Schur complement (euler_operator.h:2391-2397): CG
  // to 1e-12*|b| from a zero guess with the reference's Jacobi
  // preconditioner, fused kernels and a device-side convergence flag; the
  // host polls once per batch of iterations.
  int inner_solve(double* x, const double* b) {
    const long Ns = fine().Ns;
    const long n = 2 * Ns;
    CGState* st = cgst.p + 1;
    const double* PA = fine().inv64[BLK_A].p;
    hipLaunchKernelGGL(k_set_cg, dim3(1), dim3(1), 0, s_, st, 0.0, P.schur_inner_max, P.schur_inner_tol);
    vec_fill(n, 0.0, x, s_);
    jcg_init(Ns, fine().m, K, b, PA, in_g.p, in_d.p, st, rb, s_);
    int batch = std::max(1, last_inner);
    int i = 1, its = 0;
    while (true) {
      for (int r = 0; r < batch; ++r, ++i) {
        jcg_apply<K>(i, fine().m, rho_fix.p, PA, in_g.p, in_d.p, in_h.p, st, rb, Ns, s_);
        jcg_update(i, Ns, fine().m, K, x, in_g.p, in_d.p, in_h.p, PA, st, rb, s_);
      }
      read_state(1);
      its = hst[1].it;
      if (hst[1].done == 1) break;
      if (hst[1].done == -1) throw NoConvergence("Schur inner CG: no convergence");
      batch = 2;
    }
    last_inner = its;
    st_.inner_cg_its += its;
    return its;
  }

  // S p = D p - C A^{-1} B p (apply_add NS 2, euler_operator.h:2374-2408),
  // times the per-dof factor scale when given (GMRES's Jacobi
  // preconditioner). Plain form: B, the reference's inner CG, then D - C in
  // one pass. Mixed form: B and the cellwise-direct A^{-1} in one pass (A is
  // block diagonal; its FP64 cell inverses are rebuilt with every density).
  void schur_vmult(const double* src, double* dst, const double* scale = nullptr) {
    schur_velocity(src);
    launch_d_minus_c<K>(fine().m, oc_, bc_coeff(), pres_fix.p, s_t2.p, src, scale, dst, fine().Ns, s_);
  }

  // s_t2 = A^{-1} B src, the velocity part of the Schur product
  void schur_velocity(const double* src) {
    ++st_.schur_matvecs;
    const long Ns = fine().Ns;
    if (direct_ainv()) {
      launch_b_ainv<K>(fine().m, oc_, bc_coeff(), src, ainv.p, s_t2.p, Ns, s_);
    } else {
      launch_B<K>(fine().m, oc_, bc_coeff(), src, s_t1.p, Ns, s_);
      inner_solve(s_t2.p, s_t1.p);
    }
  }

  // The A solve the reference runs to the near-machine-precision tolerance
  // a_solve_tol (A^{-1} rhs_u, Atmospheric_Flow.cc:571): CG + MG (plain
  // form); cellwise-direct with the FP64 cell inverses (mixed form). The
  // velocity updates at the outer tolerance eps keep CG + MG in both forms.
  void solve_A_tight(double* x, const double* b) {
    if (direct_ainv()) launch_ainv_apply<K>(fine().m, ainv.p, b, x, fine().Ns, s_);
    else pcg_mg(BLK_A, x, b, P.a_solve_tol, P.max_its);
  }

  // GMRES, left Jacobi (diag D) preconditioning, convergence on the
  // preconditioned residual against tol_factor*|b| (deal.II SolverGMRES
  // defaults). Orthogonalization: classical Gram-Schmidt with a second
  // projection when the norm drops below 10*sqrt(eps) of its start (same
  // criterion as deal.II's iterated Gram-Schmidt).
  //
  // The Arnoldi vectors stay unnormalized on the device, v_i = sig_i V_i
  // with sig_i = 1/|V_i| kept in device memory, so iteration j+1 (S V_{j+1},
  // its dot products and its projection) needs nothing from the host: the
  // host keeps GM_LOOKAHEAD iterations queued and works one record behind
  // (Hessenberg column, Givens rotations, convergence), from a copy of each
  // iteration's record in pinned memory. Iterations queued past the
  // converged one are discarded; when a record asks for the second
  // projection the host drains the stream, projects again and re-queues
  // from there.
  static constexpr int GM_LOOKAHEAD = 2;

  // leading dimension of the basis: the columns are N_s doubles (a power-of-two
  // byte count on the usual meshes) plus 256 bytes apart, so the streams of
  // the multi-vector kernels do not all start on the same HBM channel
  static constexpr long GM_PAD = 32;
  long gm_ld() const { return fine().Ns + GM_PAD; }

  void gm_enqueue(int j, const double* PD) {
    const long Ns = fine().Ns;
    const long ld = gm_ld();
    double* V = gm_V.p;
    double* W = V + (long)(j + 1) * ld;
    double* rec = gm_rec.p + (long)j * GM_REC;
    // W = P S V_j with the dot products W.V_i (i <= j) and W.W of the same pass
    schur_velocity(V + (long)j * ld);
    const int G = cell_blocks<K>(fine().ncell);
    launch_d_minus_c_dots<K>(fine().m, oc_, bc_coeff(), pres_fix.p, s_t2.p, V + (long)j * ld, PD, W, Ns, V, ld, j + 1,
                             gm_part.p, G, P.arnoldi_waves, s_);
    gm_sum_slots(gm_part.p, G, G, j + 2, rec, s_);
    gm_maxpy_norm(Ns, W, V, ld, j + 1, rec, gm_sig.p, rec + GM_REC - 1, gm_sig.p + j + 1, rb, s_);
    ATMO_HIP_CHECK(hipMemcpyAsync(hgm + (long)j * GM_REC, rec, sizeof(double) * GM_REC, hipMemcpyDeviceToHost, s_));
    ATMO_HIP_CHECK(hipEventRecord(gm_ev[j], s_));
  }

  int gmres(double* x, const double* b, double tol_factor) {
    const long Ns = fine().Ns;
    const int m = P.gmres_restart;
    const double* PD = fine().inv64[BLK_D].p;
    const long ld = gm_ld();
    double* V = gm_V.p;
    double* p = gm_p.p;
    double* sig = gm_sig.p;
    double* hy = hgm + (long)MAX_GMRES * GM_REC;
    vec_dot(Ns, b, b, scal.p + 0, true, rb, s_);
    ATMO_HIP_CHECK(hipMemcpyAsync(hscal, scal.p, sizeof(double), hipMemcpyDeviceToHost, s_));
    sync();
    const double tol = tol_factor * hscal[0];
    auto check = [&](int step, double res) {
      if (res <= tol) return 1;
      if (step >= P.max_its || std::isnan(res)) return -1;
      return 0;
    };
    const double reorth_fac = 10.0 * std::sqrt(2.220446049250313e-16);
    std::vector<double> H((size_t)(m + 1) * m, 0.0), gam(m + 1), cs(m), sn(m), hs(m + 2);
    int acc = 0, state = 0;
    while (true) {
      schur_vmult(x, p);
      vec_axpby(Ns, 1.0, b, -1.0, p, s_);  // p = b - S x
      vec_mul(Ns, p, PD, V, s_);           // V_0 = P (b - S x), unnormalized
      gm_norm(Ns, V, scal.p + 0, sig, rb, s_);
      ATMO_HIP_CHECK(hipMemcpyAsync(hscal, scal.p, sizeof(double), hipMemcpyDeviceToHost, s_));
      sync();
      double rho = hscal[0];
      state = check(acc, rho);
      if (state != 0) break;
      std::fill(gam.begin(), gam.end(), 0.0);
      gam[0] = rho;
      hs[0] = 1.0 / rho;  // same correctly rounded quotient as sig_0 on the device
      int queued = 0, dim = 0;
      for (int j = 0; j < m && state == 0; ++j) {
        while (queued < m && queued - j < GM_LOOKAHEAD) gm_enqueue(queued++, PD);
        ATMO_HIP_CHECK(hipEventSynchronize(gm_ev[j]));
        ++st_.host_syncs;
        ++acc;
        dim = j + 1;
        const double* r = hgm + (long)j * GM_REC;
        double* hc = &H[(size_t)j * (m + 1)];
        for (int i = 0; i <= j; ++i) hc[i] = hs[j] * hs[i] * r[i];
        double nraw = r[GM_REC - 1];
        if (nraw < reorth_fac * std::sqrt(r[j + 1])) {
          // second projection of W = V_{j+1}; iterations queued after j used the first one
          sync();
          double* W = V + (long)(j + 1) * ld;
          double* c2 = gm_coef.p;
          gm_mdot(Ns, W, V, ld, j + 1, c2, rb, s_);
          gm_maxpy_norm(Ns, W, V, ld, j + 1, c2, sig, scal.p + 1, sig + j + 1, rb, s_);
          ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 2, c2, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, s_));
          ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 1, scal.p + 1, sizeof(double), hipMemcpyDeviceToHost, s_));
          sync();
          for (int i = 0; i <= j; ++i) hc[i] += hs[j] * hs[i] * hscal[2 + i];
          nraw = hscal[1];
          queued = j + 1;
        }
        hs[j + 1] = nraw > 0.0 ? 1.0 / nraw : 0.0;
        hc[j + 1] = hs[j] * nraw;  // |w| of the actual (normalized-basis) Arnoldi vector
        for (int i = 0; i < j; ++i) {
          const double t = hc[i];
          hc[i] = cs[i] * t + sn[i] * hc[i + 1];
          hc[i + 1] = -sn[i] * t + cs[i] * hc[i + 1];
        }
        const double rr = 1.0 / std::sqrt(hc[j] * hc[j] + hc[j + 1] * hc[j + 1]);
        sn[j] = hc[j + 1] * rr;
        cs[j] = hc[j] * rr;
        hc[j] = cs[j] * hc[j] + sn[j] * hc[j + 1];
        gam[j + 1] = -sn[j] * gam[j];
        gam[j] *= cs[j];
        rho = std::abs(gam[j + 1]);
        state = check(acc, rho);
      }
      std::vector<double> y(dim);
      for (int i = dim - 1; i >= 0; --i) {
        double sacc = gam[i];
        for (int l = i + 1; l < dim; ++l) sacc -= H[(size_t)l * (m + 1) + i] * y[l];
        y[i] = sacc / H[(size_t)i * (m + 1) + i];
      }
      // x += sum_i y_i v_i = sum_i (y_i sig_i) V_i; the pinned coefficients are
      // not rewritten before the next drain of the stream
      for (int i = 0; i < dim; ++i) hy[i] = y[i] * hs[i];
      ATMO_HIP_CHECK(hipMemcpyAsync(gm_coef.p, hy, sizeof(double) * dim, hipMemcpyHostToDevice, s_));
      gm_update(Ns, x, V, ld, dim, gm_coef.p, s_);
      if (state != 0) break;
    }
    if (state == -1) {
      sync();
      throw NoConvergence("GMRES: no convergence");
    }
    st_.gmres_its += acc;
    return acc;
  }

  // ---------------------------------------------------- solver stages
  // update_density (Atmospheric_Flow.cc:430-537)
  void update_density(int hs) {
    set_stages(hs, hs == 3 ? 4 : 1);
    const long Ns = fine().Ns;
    double* rhs = gm_p.p;  // scratch of size Ns
    if (hs == 1) {
      const double* s[2] = {rho_old.p, u_old.p};
      rhs_rho(s, 2, rhs);
      vec_copy(Ns, rho_old.p, rho2.p, s_);
      pcg_mg(BLK_RHO, rho2.p, rhs, P.eps, P.max_its);
    } else if (hs == 2) {
      const double* s[4] = {rho_old.p, u_old.p, rho2.p, u2.p};
      rhs_rho(s, 4, rhs);
      vec_copy(Ns, rho2.p, rho3.p, s_);
      pcg_mg(BLK_RHO, rho3.p, rhs, P.eps, P.max_its);
    } else {
      const double* s[6] = {rho_old.p, u_old.p, rho2.p, u2.p, rho3.p, u3.p};
      rhs_rho(s, 6, rhs);
      vec_copy(Ns, rho3.p, rho_curr.p, s_);
      pcg_mg(BLK_RHO, rho_curr.p, rhs, P.eps, P.max_its);
    }
  }

  // pressure_fixed_point (Atmospheric_Flow.cc:544-673)
  // rhs_u and tmp_1 = A^{-1} rhs_u do not depend on the fixed-point iterate: formed in the stage's first
  // iteration only (the reference recomputes both every iteration, Atmospheric_Flow.cc:558-571; its A solve
  // then restarts from the previous tmp_1, whose residual already meets the tolerance, and returns it
  // unchanged, so the values are the same)
  void pressure_fixed_point(int hs, bool first) {
    set_stages(hs, 2);
    set_pres_fixed(p_fixed_old.p);
    const long Ns = fine().Ns;
    if (hs == 1) {
      const double* sp[5] = {rho_old.p, u_old.p, p_old.p, rho2.p, u_fixed.p};
      rhs_pressure(sp, 5, rhs_p.p);
      const double* su[4] = {rho_old.p, u_old.p, p_old.p, rho2.p};
      if (first) rhs_velocity(su, 4, rhs_u0.p);
    } else {
      const double* sp[8] = {rho_old.p, u_old.p, p_old.p, rho2.p, u2.p, p2.p, rho3.p, u_fixed.p};
      rhs_pressure(sp, 8, rhs_p.p);
      const double* su[7] = {rho_old.p, u_old.p, p_old.p, rho2.p, u2.p, p2.p, rho3.p};
      if (first) rhs_velocity(su, 7, rhs_u0.p);
    }
    if (first) solve_A_tight(tmp1.p, rhs_u0.p);
    vec_copy(2 * Ns, rhs_u0.p, rhs_u.p, s_);
    vmult_enthalpy(tmp1.p, s_t3.p);
    vec_axpby(Ns, -1.0, s_t3.p, 1.0, rhs_p.p, s_);
    vec_copy(Ns, p_fixed_old.p, p_fixed.p, s_);
    gmres(p_fixed.p, rhs_p.p, P.eps);
  }

  // update_velocity (Atmospheric_Flow.cc:680-781)
  void update_velocity(int hs) {
    const long Ns = fine().Ns;
    if (hs == 1 || hs == 2) {
      set_stages(hs, 3);
      vmult_pressure(p_fixed.p, s_t1.p);
      vec_axpby(2 * Ns, -1.0, s_t1.p, 1.0, rhs_u.p, s_);
      pcg_mg(BLK_A, u_fixed.p, rhs_u.p, P.eps, P.max_its);
    } else {
      set_stages(3, 5);
      const double* su[9] = {rho_old.p, u_old.p, p_old.p, rho2.p, u2.p, p2.p, rho3.p, u3.p, p3.p};
      rhs_velocity(su, 9, rhs_u.p);
      vec_copy(2 * Ns, u3.p, u_curr.p, s_);
      pcg_mg(BLK_A, u_curr.p, rhs_u.p, P.eps, P.max_its);
    }
  }

  // update_pressure (Atmospheric_Flow.cc:788-871)
  void update_pressure() {
    set_stages(3, 6);
    const double* sp[11] = {rho_old.p, u_old.p, p_old.p, rho2.p, u2.p, p2.p, rho3.p, u3.p, p3.p, rho_curr.p, u_curr.p};
    rhs_pressure(sp, 11, rhs_p.p);
    vec_copy(fine().Ns, p3.p, p_old.p, s_);
    pcg_mg(BLK_D, p_old.p, rhs_p.p, P.eps, P.max_its);
  }

  // fixed-point loop with the relative L-infinity stop test on nodal
  // pressure values (Atmospheric_Flow.cc:1135-1153)
  void fixed_point_loop(int hs) {
    const long Ns = fine().Ns;
    for (int iter = 0; iter < P.fp_max_iter; ++iter) {
      pressure_fixed_point(hs, iter == 0);
      update_velocity(hs);
      ++st_.fp_its;
      vec_linf(Ns, p_fixed.p, nullptr, scal.p + 2, rb, s_);
      vec_linf(Ns, p_fixed.p, p_fixed_old.p, scal.p + 3, rb, s_);
      ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 2, scal.p + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, s_));
      sync();
      const double err = hscal[3] / (hscal[2] + 1e-10);
      if (P.fp_lock > 0) {
        if (iter + 1 >= P.fp_lock) break;
      } else if (err < P.fp_tol) {
        break;
      }
      vec_copy(Ns, p_fixed.p, p_fixed_old.p, s_);
    }
  }
};

template <int K>
template <typename T>
void EngineK<K>::level_apply(Level& L, int blk, const T* src, T* dst, T cf, const T* add, T add_scale,
                             hipStream_t s) {
  const T* rho = nullptr;
  if (blk == BLK_A) rho = std::is_same<T, double>::value ? (const T*)L.rho64.p : (const T*)L.rho32.p;
  launch_mass<K, T>(L.m, ncomp(blk), blk == BLK_A, src, rho, dst, (T)(cf * coeff(blk)), add, add_scale, L.Ns,
                    s ? s : s_);
}

// Mirrors the allocations of EngineK (constructor, build_levels): device
// bytes an engine of this mesh and parameters requests.
size_t engine_device_bytes(int k, const MeshDesc& mesh, const EngineParams& p) {
  if (k < 1 || k > 3 || mesh.nx < 1 || mesh.nz < 1) throw std::invalid_argument("engine_device_bytes: bad mesh");
  const size_t nd2 = (size_t)(k + 1) * (k + 1), D = sizeof(double), F = sizeof(float);
  size_t b = 0;
  int nx = mesh.nx, nz = mesh.nz, l = 0;
  std::vector<std::pair<int, int>> lev{{nx, nz}};
  while (nx % 2 == 0 && nz % 2 == 0 && nx >= 4 && (nx * nz) / 4 >= p.mg_min_cells) {
    nx /= 2;
    nz /= 2;
    lev.push_back({nx, nz});
  }
  for (auto it = lev.rbegin(); it != lev.rend(); ++it, ++l) {
    const size_t Ns = (size_t)it->first * it->second * nd2;
    b += (it->first + 1) * D;             // h
    b += Ns * D + Ns * F;                 // rho64, rho32
    b += 4 * Ns * D + 4 * Ns * F;         // inv64[3], inv32[3]
    b += 8 * Ns * D + 8 * Ns * F;         // x,b,t,d in both precisions
    if (l == 0) b += 6 * Ns * D;          // coarse CG work
    b += 3 * Ns * D;                      // eigenvalue start vectors
  }
  const size_t Ns = (size_t)mesh.nx * mesh.nz * nd2;
  b += RedBuf::doubles() * D + 4 * sizeof(CGState) + 256 * D + sizeof(int) + DIAG_PARTIALS * D + DG_SLOTS * D;
  b += lev.size() * (2 * MAX_EIG_ITS * D + sizeof(CGState));  // eigenvalue estimates: records and CG states
  b += ENGINE_EIG_STREAMS * RED_MAX_BLOCKS * D;               // their reduction partials on the extra streams
  b += 2 * Ns * D;                        // rho_fix, pres_fix
  b += 8 * 2 * Ns * D + Ns * D;           // CG / Schur / inner work
  if (p.precision >= 1) b += (size_t)mesh.nx * mesh.nz * nd2 * (nd2 + 1) / 2 * D;  // packed FP64 cell-block inverses of A
  b += (size_t)(p.gmres_restart + 2) * (Ns + 32) * D + Ns * D + (MAX_GMRES + 2) * D;  // basis: leading dim N_s + 32
  b += (size_t)MAX_GMRES * GM_REC * D + (MAX_GMRES + 2) * D;  // GMRES records and basis scales
  b += (size_t)RED_MAX_SLOTS * ((mesh.nx * mesh.nz + 64 / (2 * k + 1) - 1) / (64 / (2 * k + 1))) * D;  // Arnoldi partials
  b += 10 * Ns * D + 8 * 2 * Ns * D;      // IMEX state and stage fields
  return b;
}

std::unique_ptr<Engine> Engine::create_amr(int k, const MeshDesc& base, const std::vector<QCell>& leaves,
                                           const EngineParams& p) {
  switch (k) {
    case 1: return std::make_unique<EngineK<1>>(base, leaves, p);
    case 2: return std::make_unique<EngineK<2>>(base, leaves, p);
    case 3: return std::make_unique<EngineK<3>>(base, leaves, p);
    default: throw std::invalid_argument("degree must be 1, 2 or 3");
  }
}

std::unique_ptr<Engine> Engine::create(int k, const MeshDesc& mesh, const EngineParams& p) {
  switch (k) {
    case 1: return std::make_unique<EngineK<1>>(mesh, p);
    case 2: return std::make_unique<EngineK<2>>(mesh, p);
    case 3: return std::make_unique<EngineK<3>>(mesh, p);
    default: throw std::invalid_argument("degree must be 1, 2 or 3");
  }
}

// ------------------------------------------------------- initial state
static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Hydrostatic background with constant buoyancy frequency, interpolated at
// the FE_DGQ support points (VectorTools::interpolate, Atmospheric_Flow.cc:420-422),
// density perturbed by 1 + pert*(2U-1), U = top 53 bits of
// splitmix64(seed ^ global dof id) / 2^53.
// Hill state on cells given by their four vertex heights zc[4*c..] (x plays no role).
static void hill_state_cells(int k, long ncell, const double* zc, double theta0, double p0, double Nbv, double U,
                             double g, double Rgas, double gamma_gas, double pert, uint64_t seed, double* rho,
                             double* u, double* p) {
  constexpr int KMAX = 3;
  if (k < 1 || k > KMAX) throw std::invalid_argument("degree must be 1..3");
  double gll[KMAX + 1];
  if (k == 1) { gll[0] = 0; gll[1] = 1; }
  else if (k == 2) { for (int i = 0; i < 3; ++i) gll[i] = basis::kT<2>.gll[i]; }
  else { for (int i = 0; i < 4; ++i) gll[i] = basis::kT<3>.gll[i]; }
  const int nd = k + 1, nd2 = nd * nd;
  const long Ns = ncell * nd2;
  const double cp = gamma_gas * Rgas / (gamma_gas - 1.0);
  for (long c = 0; c < ncell; ++c) {
    const double* z = zc + 4 * c;
    for (int jj = 0; jj < nd; ++jj)
      for (int ii = 0; ii < nd; ++ii) {
        const double xi = gll[ii], eta = gll[jj];
        const double zz = (1 - xi) * (1 - eta) * z[0] + xi * (1 - eta) * z[1] + (1 - xi) * eta * z[2] + xi * eta * z[3];
        const double th = theta0 * std::exp(Nbv * Nbv * zz / g);
        const double ex = 1.0 + g * g / (cp * theta0 * Nbv * Nbv) * (std::exp(-Nbv * Nbv * zz / g) - 1.0);
        const double pr = p0 * std::pow(ex, cp / Rgas);
        const double rh = pr / (Rgas * th * ex);
        const long id = (long)c * nd2 + jj * nd + ii;
        const uint64_t r = splitmix64(seed ^ (uint64_t)id);
        const double U01 = (double)(r >> 11) * (1.0 / 9007199254740992.0);
        rho[id] = rh * (1.0 + pert * (2.0 * U01 - 1.0));
        p[id] = pr;
        u[id] = U;
        u[Ns + id] = 0.0;
      }
  }
}

void hill_initial_state(int k, const MeshDesc& m, double theta0, double p0, double Nbv, double U, double g,
                        double Rgas, double gamma_gas, double pert, uint64_t seed, double* rho, double* u,
                        double* p) {
  std::vector<double> zc(4 * (size_t)m.nx * m.nz);
  for (int c = 0; c < m.nx * m.nz; ++c) {
    const int i = c % m.nx, j = c / m.nx;
    const double hl = m.h[i], hr = m.h[i + 1];
    const double z[4] = {hl + (m.H - hl) * (double)j / m.nz, hr + (m.H - hr) * (double)j / m.nz,
                         hl + (m.H - hl) * (double)(j + 1) / m.nz, hr + (m.H - hr) * (double)(j + 1) / m.nz};
    for (int a = 0; a < 4; ++a) zc[4 * c + a] = z[a];
  }
  hill_state_cells(k, (long)m.nx * m.nz, zc.data(), theta0, p0, Nbv, U, g, Rgas, gamma_gas, pert, seed, rho, u, p);
}

void hill_initial_state_cells(int k, const std::vector<double>& cgeo, double theta0, double p0, double Nbv, double U,
                              double g, double Rgas, double gamma_gas, double pert, uint64_t seed, double* rho,
                              double* u, double* p) {
  const long nc = (long)cgeo.size() / 6;
  std::vector<double> zc(4 * nc);
  for (long c = 0; c < nc; ++c)
    for (int a = 0; a < 4; ++a) zc[4 * c + a] = cgeo[6 * c + 2 + a];
  hill_state_cells(k, nc, zc.data(), theta0, p0, Nbv, U, g, Rgas, gamma_gas, pert, seed, rho, u, p);
}

}  // namespace atmo
