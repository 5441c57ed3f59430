// Device engine of the spherical-shell path (shell.h): the EULEROperator
// interface (euler_operator.h:24-211) and the EulerSolver stage sequence
// (Atmospheric_Flow.cc:430-871, 1100-1257) in dim = 3 at FE_DGQ(1), every
// field resident in HBM, on one HIP stream.
//
// Solvers. M_rho, A and D are block diagonal in DG (one 8 x 8 block per cell
// and component), so every mass-type solve of the stage sequence (the
// reference's CG + MG, Atmospheric_Flow.cc:452-536, 701-870, and the Schur
// complement's inner CG, euler_operator.h:2383-2400) is the cell-by-cell
// product with the blocks' FP64 inverses: the exact solution the reference's
// iterations approach. M's inverse blocks are built once, A's whenever the
// density of the fixed point changes. The Schur complement is solved by
// GMRES with left Jacobi preconditioning (diag D) and deal.II's stopping
// rule, its Arnoldi vectors unnormalized on the device as on the terrain
// path. Every precision setting runs this FP64 form (the shell has no
// multigrid hierarchy).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "dbuf.h"
#include "diag.h"
#include "engine.h"
#include "krylov.h"
#include "ops.h"
#include "shell.h"

namespace atmo {

namespace {

constexpr int GM_HOST_S = (MAX_GMRES + 1) * GM_REC + MAX_GMRES + 2;

class ShellEngine : public Engine {
 public:
  ShellEngine(const ShellDesc& d, const EngineParams& p) : P(p) {
    const size_t bytes0 = tl_dev_bytes;
    mesh_ = build_shell(d);
    ATMO_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    const long Ns = ns();
    const int nc = mesh_.ncell;
    corners_.alloc(mesh_.corners.size());
    nb_.alloc(mesh_.nb.size());
    nbc_.alloc(mesh_.nbc.size());
    ATMO_HIP_CHECK(hipMemcpyAsync(corners_.p, mesh_.corners.data(), corners_.n * sizeof(double), hipMemcpyHostToDevice, s_));
    ATMO_HIP_CHECK(hipMemcpyAsync(nb_.p, mesh_.nb.data(), nb_.n * sizeof(int), hipMemcpyHostToDevice, s_));
    ATMO_HIP_CHECK(hipMemcpyAsync(nbc_.p, mesh_.nbc.data(), nbc_.n, hipMemcpyHostToDevice, s_));
    m_.n = mesh_.n;
    m_.ncell = nc;
    m_.r_in = d.r_inner;
    m_.dr = (d.r_outer - d.r_inner) / mesh_.n;
    m_.planet2 = d.planet * d.planet;
    m_.corners = corners_.p;
    m_.nb = nb_.p;
    m_.nbc = nbc_.p;
    min_diam_ = shell_min_cell_diameter(mesh_);

    red_partials.alloc(RedBuf::doubles());
    red_partials.zero(s_);
    rb.bind(red_partials.p);
    scal.alloc(256);
    scal.zero(s_);
    diag_part.alloc(DIAG_PARTIALS);
    diag_rec.alloc(DG_SLOTS);
    diag_rec.zero(s_);
    ATMO_HIP_CHECK(hipHostMalloc((void**)&hscal, sizeof(double) * 256, hipHostMallocDefault));
    ATMO_HIP_CHECK(hipHostMalloc((void**)&hgm, sizeof(double) * GM_HOST_S, hipHostMallocDefault));
    for (auto& e : gm_ev) ATMO_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    minv.alloc((size_t)nc * 36);
    mdiag.alloc((size_t)nc * 8);
    ainv.alloc((size_t)nc * 36);
    adiag.alloc((size_t)nc * 8);
    ddiag.alloc(Ns);
    rho_fix.alloc(Ns);
    pres_fix.alloc(Ns);
    for (auto* b : {&rho_fix, &pres_fix}) b->zero(s_);
    for (auto* b : {&s_t1, &s_t2}) b->alloc(3 * Ns);
    s_t3.alloc(Ns);
    gm_V.alloc((size_t)(P.gmres_restart + 2) * gm_ld());
    gm_p.alloc(Ns);
    gm_coef.alloc(MAX_GMRES + 2);
    gm_rec.alloc((size_t)MAX_GMRES * GM_REC);
    gm_sig.alloc(MAX_GMRES + 2);
    gm_part.alloc((size_t)RED_MAX_SLOTS * shell_blocks(m_.ncell));
    for (auto* b : {&rho_old, &rho2, &rho3, &rho_curr, &p_old, &p2, &p3, &p_fixed, &p_fixed_old, &rhs_p}) b->alloc(Ns);
    for (auto* b : {&u_old, &u2, &u3, &u_curr, &u_fixed, &tmp1, &rhs_u, &rhs_u0}) b->alloc(3 * Ns);
    for (auto* b : {&rho_old, &rho2, &rho3, &rho_curr, &p_old, &p2, &p3, &p_fixed, &p_fixed_old, &rhs_p, &u_old, &u2,
                    &u3, &u_curr, &u_fixed, &tmp1, &rhs_u, &rhs_u0})
      b->zero(s_);
    // geometry-only blocks: M^{-1} and diag(D)^{-1} = (gamma - 1) diag(M)^{-1}
    shell_block_inverse(m_, nullptr, minv.p, mdiag.p, nullptr, s_);
    vec_fill(Ns, 0.0, ddiag.p, s_);
    vec_axpby(Ns, P.gamma_gas - 1.0, mdiag.p, 0.0, ddiag.p, s_);
    oc_ = consts();
    sync();
    bytes_ = tl_dev_bytes - bytes0;
  }

  ~ShellEngine() override {
    (void)hipStreamSynchronize(s_);
    if (hscal) (void)hipHostFree(hscal);
    if (hgm) (void)hipHostFree(hgm);
    for (auto& e : gm_ev)
      if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(s_);
  }

  int dim() const override { return 3; }
  int degree() const override { return 1; }
  int ncells() const override { return mesh_.ncell; }
  long ns() const override { return (long)mesh_.ncell * 8; }
  int nlevels() const override { return 1; }
  hipStream_t stream() const override { return s_; }
  const EngineParams& params() const override { return P; }
  EngineStats& stats() override { return st_; }
  size_t device_bytes() const override { return bytes_; }
  const std::vector<double>& cell_geometry() const override { return no_cells_; }
  double min_cell_diameter() const override { return min_diam_; }
  const ShellMesh* shell() const override { return &mesh_; }

  void set_params(const EngineParams& p) override {
    if (p.gmres_restart > P.gmres_restart) gm_V.alloc((size_t)(p.gmres_restart + 2) * gm_ld());
    const bool new_gamma = p.gamma_gas != P.gamma_gas;
    P = p;
    oc_ = consts();
    if (new_gamma) vec_axpby(ns(), P.gamma_gas - 1.0, mdiag.p, 0.0, ddiag.p, s_);
  }

  void set_stages(int hs, int ns_) override {
    if (hs < 1 || hs > 3) throw std::invalid_argument("HYPERBOLIC_stage must be 1, 2 or 3");
    if (ns_ < 1 || ns_ > 6) throw std::invalid_argument("NS_stage must be in [1, 6]");
    hyper = hs;
    nsst = ns_;
  }
  // also the A blocks' inverses; a block that is not positive definite (the
  // density is not positive inside a cell) is flagged on the device and
  // reported by the next solve that fails
  void set_rho_for_fixed(const double* rho) override {
    vec_copy(ns(), rho, rho_fix.p, s_);
    vec_fill(1, 0.0, scal.p + 250, s_);
    shell_block_inverse(m_, rho_fix.p, ainv.p, adiag.p, scal.p + 250, s_);
  }
  void set_pres_fixed(const double* p) override { vec_copy(ns(), p, pres_fix.p, s_); }

  void rhs_rho(const double* const* src, int nsrc, double* dst) override {
    if (nsrc != 2 * hyper) throw std::invalid_argument("vmult_rhs_rho_update: expected 2*stage sources");
    shell_rhs_rho(m_, oc_, hyper, src, dst, s_);
  }
  void rhs_pressure(const double* const* src, int nsrc, double* dst) override {
    const int need = hyper == 1 ? 5 : (hyper == 2 ? 8 : 11);
    if (nsrc != need) throw std::invalid_argument("vmult_rhs_pressure: wrong number of sources");
    shell_rhs_pressure(m_, oc_, hyper, src, dst, s_);
  }
  void rhs_velocity(const double* const* src, int nsrc, double* dst) override {
    const int need = hyper == 1 ? 4 : (hyper == 2 ? 7 : 9);
    if (nsrc != need) throw std::invalid_argument("vmult_rhs_velocity_fixed: wrong number of sources");
    shell_rhs_velocity(m_, oc_, hyper, src, dst, s_);
  }
  double bc_coeff() const { return hyper == 1 ? tb_.a22t : tb_.a33t; }
  void vmult_pressure(const double* p, double* dst_u) override { shell_B(m_, oc_, bc_coeff(), p, dst_u, s_); }
  void vmult_enthalpy(const double* u, double* dst_p) override {
    shell_C(m_, oc_, bc_coeff(), pres_fix.p, u, dst_p, s_);
  }

  void vmult(const double* src, double* dst) override {
    if (nsst == 2) schur_vmult(src, dst);
    else if (nsst == 3 || nsst == 5) shell_mass(m_, 3, rho_fix.p, 1.0, src, dst, s_);
    else shell_mass(m_, 1, nullptr, nsst == 6 ? 1.0 / (P.gamma_gas - 1.0) : 1.0, src, dst, s_);
  }

  void inverse_diagonal(double* dst) override {
    const long Ns = ns();
    if (nsst == 3 || nsst == 5) {
      for (int c = 0; c < 3; ++c) vec_copy(Ns, adiag.p, dst + c * Ns, s_);
    } else {
      vec_copy(Ns, (nsst == 2 || nsst == 6) ? ddiag.p : mdiag.p, dst, s_);
    }
  }

  // The mass-type block of ns (1/4 M_rho, 3/5 A, 6 D), solved cell by cell;
  // tol_factor is met by the exact solve. Returns 0 iterations.
  int solve_mass_mg(int ns_, double* x, const double* b, double tol_factor) override {
    (void)tol_factor;
    if (ns_ == 3 || ns_ == 5) shell_block_apply(m_, 3, ainv.p, 1.0, b, x, s_);
    else if (ns_ == 6) shell_block_apply(m_, 1, minv.p, P.gamma_gas - 1.0, b, x, s_);
    else if (ns_ == 1 || ns_ == 4) shell_block_apply(m_, 1, minv.p, 1.0, b, x, s_);
    else throw std::invalid_argument("ns_stage must be 1, 3, 4, 5 or 6");
    return 0;
  }
  int solve_schur(double* x, const double* b, double tol_factor) override { return gmres(x, b, tol_factor); }

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

  void load_state(const double* rho, const double* u, const double* p) override {
    const long Ns = ns();
    vec_copy(Ns, rho, rho_old.p, s_);
    vec_copy(3 * Ns, u, u_old.p, s_);
    vec_copy(Ns, p, p_old.p, s_);
    tmp1.zero(s_);
  }
  void store_state(double* rho, double* u, double* p) override {
    const long Ns = ns();
    vec_copy(Ns, rho_old.p, rho, s_);
    vec_copy(3 * Ns, u_old.p, u, s_);
    vec_copy(Ns, p_old.p, p, s_);
  }
  double* state_ptr(int which) override { return which == 0 ? rho_old.p : (which == 1 ? u_old.p : p_old.p); }

  // one IMEX step (Atmospheric_Flow.cc:1115-1222)
  void step() override {
    const long Ns = ns();
    update_density(1);
    launch_diag_density_n(Ns, rho2.p, 0, diag_part.p, diag_rec.p, s_);
    set_rho_for_fixed(rho2.p);
    vec_copy(Ns, p_old.p, p_fixed_old.p, s_);
    vec_copy(3 * Ns, u_old.p, u_fixed.p, s_);
    fixed_point_loop(1);
    vec_copy(Ns, p_fixed.p, p2.p, s_);
    vec_copy(3 * Ns, u_fixed.p, u2.p, s_);
    update_density(2);
    launch_diag_density_n(Ns, rho3.p, 1, diag_part.p, diag_rec.p, s_);
    set_rho_for_fixed(rho3.p);
    vec_copy(Ns, p2.p, p_fixed_old.p, s_);
    vec_copy(3 * Ns, u2.p, u_fixed.p, s_);
    fixed_point_loop(2);
    vec_copy(Ns, p_fixed.p, p3.p, s_);
    vec_copy(3 * Ns, u_fixed.p, u3.p, s_);
    update_density(3);
    launch_diag_density_n(Ns, rho_curr.p, 2, diag_part.p, diag_rec.p, s_);
    set_rho_for_fixed(rho_curr.p);
    update_velocity(3);
    update_pressure();
    vec_copy(Ns, rho_curr.p, rho_old.p, s_);
    vec_copy(3 * Ns, u_curr.p, u_old.p, s_);
    launch_diag_flow<1>(MeshDev{}, rho_old.p, u_old.p, p_old.p, P.gamma_gas, Ns, diag_part.p, diag_rec.p, s_);
    ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 250, scal.p + 250, sizeof(double), hipMemcpyDeviceToHost, s_));
    sync();
    if (hscal[250] != 0.0) throw_bad_density();
  }

  void step_diagnostics(double* rec, bool of_state) override {
    const long Ns = ns();
    if (of_state) {
      launch_diag_density_n(Ns, rho_old.p, 2, diag_part.p, diag_rec.p, s_);
      launch_diag_flow<1>(MeshDev{}, rho_old.p, u_old.p, p_old.p, P.gamma_gas, Ns, diag_part.p, diag_rec.p, s_);
    }
    ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 128, diag_rec.p, sizeof(double) * DG_SLOTS, hipMemcpyDeviceToHost, s_));
    sync();
    std::memcpy(rec, hscal + 128, sizeof(double) * DG_SLOTS);
  }

  // which: 2 M apply, 3 B, 4 C, 5 stage-3 pressure residual, 7 device copy of
  // one velocity field (calibration), 9 scaled D - C, 14 stage-3 velocity
  // residual, 15 stage-3 density residual
  void time_kernel(int which, int nrep, double* ms, double* bytes) override {
    const long Ns = ns();
    double fields = 0;
    hipEvent_t e0, e1;
    ATMO_HIP_CHECK(hipEventCreate(&e0));
    ATMO_HIP_CHECK(hipEventCreate(&e1));
    const double* s11[11] = {rho_old.p, u_old.p, p_old.p, rho_old.p, u_old.p, p_old.p,
                             rho_old.p, u_old.p, p_old.p, rho_old.p, u_old.p};
    const double* s6[6] = {rho_old.p, u_old.p, rho_old.p, u_old.p, rho_old.p, u_old.p};
    auto launch = [&]() {
      switch (which) {
        case 2:
          shell_mass(m_, 1, nullptr, 1.0, rho_old.p, gm_p.p, s_);
          fields = 2;
          break;
        case 3:
          shell_B(m_, oc_, bc_coeff(), p_old.p, s_t1.p, s_);
          fields = 4;
          break;
        case 4:
          shell_C(m_, oc_, bc_coeff(), pres_fix.p, u_old.p, gm_p.p, s_);
          fields = 5;
          break;
        case 5:
          shell_rhs_pressure(m_, oc_, 3, s11, gm_p.p, s_);
          fields = 20;
          break;
        case 7:
          ATMO_HIP_CHECK(hipMemcpyAsync(s_t1.p, u_old.p, 3 * Ns * sizeof(double), hipMemcpyDeviceToDevice, s_));
          fields = 6;
          break;
        case 9:
          shell_d_minus_c(m_, oc_, bc_coeff(), pres_fix.p, u_old.p, p_old.p, ddiag.p, gm_p.p, s_);
          fields = 7;
          break;
        case 8:  // B with the cellwise A^{-1} in one pass: p in, y (3 components) out, the packed blocks
          shell_b_ainv(m_, oc_, bc_coeff(), p_old.p, ainv.p, s_t2.p, s_);
          fields = 4 + 36.0 / 8.0;
          break;
        case 12: {  // the Arnoldi step's scaled D - C with the dot products against 14 basis vectors
          if (P.gmres_restart < 13) throw std::invalid_argument("time_kernel 12 needs gmres_restart >= 13");
          // as GMRES runs it: applied to the newest basis vector V_13, whose staged values serve its own
          // dot product (p, y (3 components), p_fix, the Jacobi scale and 13 other basis vectors in; w out)
          const long ld = gm_ld();
          const int G = shell_blocks(m_.ncell);
          shell_d_minus_c_dots(m_, oc_, bc_coeff(), pres_fix.p, u_old.p, gm_V.p + 13 * ld, ddiag.p,
                               gm_V.p + 14 * ld, gm_V.p, ld, 14, gm_part.p, G, s_);
          fields = 7 + 13;
          break;
        }
        case 14:
          shell_rhs_velocity(m_, oc_, 3, s11, s_t1.p, s_);
          fields = 18;
          break;
        case 15:
          shell_rhs_rho(m_, oc_, 3, s6, gm_p.p, s_);
          fields = 13;
          break;
        default:
          throw std::invalid_argument("time_kernel: unknown kernel id for the shell");
      }
    };
    launch();
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

 private:
  EngineParams P;
  Tableau tb_;
  ShellConsts oc_{};
  ShellMesh mesh_;
  ShellDev m_{};
  double min_diam_ = 0;
  std::vector<double> no_cells_;
  hipStream_t s_ = nullptr;
  EngineStats st_;
  size_t bytes_ = 0;
  int hyper = 1, nsst = 1;

  DBuf<double> corners_;
  DBuf<int> nb_;
  DBuf<unsigned char> nbc_;
  DBuf<double> red_partials, scal, diag_part, diag_rec;
  RedBuf rb;
  double* hscal = nullptr;
  double* hgm = nullptr;
  hipEvent_t gm_ev[MAX_GMRES] = {};
  DBuf<double> minv, mdiag, ainv, adiag, ddiag, rho_fix, pres_fix;
  DBuf<double> s_t1, s_t2, s_t3;
  DBuf<double> gm_V, gm_p, gm_coef, gm_rec, gm_sig;
  DBuf<double> gm_part;  // per-block partial dot products of the fused Arnoldi step, [slot][block]
  DBuf<double> rho_old, rho2, rho3, rho_curr, u_old, u2, u3, u_curr, u_fixed, p_old, p2, p3, p_fixed, p_fixed_old,
      tmp1, rhs_u, rhs_p, rhs_u0;

  ShellConsts consts() const { return ShellConsts{P.Ma, P.Fr, P.dt, P.gamma_gas}; }

  // A's cell blocks were not all positive definite: the fixed point's density
  // is not positive everywhere (the reference's CG on A would not converge)
  [[noreturn]] static void throw_bad_density() {
    throw std::runtime_error("non-finite A^{-1}: the density of the fixed point is not positive in every cell "
                             "(A is not positive definite)");
  }
  void sync() { ATMO_HIP_CHECK(hipStreamSynchronize(s_)); }

  // S p = D p - C A^{-1} B p (apply_add NS 2, euler_operator.h:2374-2408), times scale when given
  void schur_vmult(const double* src, double* dst, const double* scale = nullptr) {
    ++st_.schur_matvecs;
    shell_b_ainv(m_, oc_, bc_coeff(), src, ainv.p, s_t2.p, s_);
    shell_d_minus_c(m_, oc_, bc_coeff(), pres_fix.p, s_t2.p, src, scale, dst, s_);
  }

  // GMRES as on the terrain path: left Jacobi (diag D), convergence of the
  // preconditioned residual against tol_factor*|b| (deal.II SolverGMRES),
  // classical Gram-Schmidt with a second projection when the norm drops
  // below 10*sqrt(eps) of its start; basis vectors unnormalized on the
  // device (v_i = sig_i V_i), the host a record behind with GM_LOOKAHEAD
  // iterations queued.
  static constexpr int GM_LOOKAHEAD = 2;
  static constexpr long GM_PAD = 32;
  long gm_ld() const { return ns() + GM_PAD; }

  void gm_enqueue(int j) {
    const long Ns = ns(), ld = gm_ld();
    double* V = gm_V.p;
    double* W = V + (long)(j + 1) * ld;
    double* rec = gm_rec.p + (long)j * GM_REC;
    // W = P S V_j (the Schur product: A^{-1} B in one pass, then D - C) with the dot products W.V_i (i <= j)
    // and W.W formed in the D - C pass
    const double* Vj = V + (long)j * ld;
    ++st_.schur_matvecs;
    shell_b_ainv(m_, oc_, bc_coeff(), Vj, ainv.p, s_t2.p, s_);
    const int G = shell_blocks(m_.ncell);
    shell_d_minus_c_dots(m_, oc_, bc_coeff(), pres_fix.p, s_t2.p, Vj, ddiag.p, W, V, ld, j + 1, gm_part.p, G, s_);
    gm_sum_slots(gm_part.p, G, G, j + 2, rec, s_);
    gm_maxpy_norm(Ns, W, V, ld, j + 1, rec, gm_sig.p, rec + GM_REC - 1, gm_sig.p + j + 1, rb, s_);
    ATMO_HIP_CHECK(hipMemcpyAsync(hgm + (long)j * GM_REC, rec, sizeof(double) * GM_REC, hipMemcpyDeviceToHost, s_));
    ATMO_HIP_CHECK(hipEventRecord(gm_ev[j], s_));
  }

  int gmres(double* x, const double* b, double tol_factor) {
    const long Ns = ns();
    const int m = P.gmres_restart;
    const double* PD = ddiag.p;
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
      hs[0] = 1.0 / rho;
      int queued = 0, dim = 0;
      for (int j = 0; j < m && state == 0; ++j) {
        while (queued < m && queued - j < GM_LOOKAHEAD) gm_enqueue(queued++);
        ATMO_HIP_CHECK(hipEventSynchronize(gm_ev[j]));
        ++st_.host_syncs;
        ++acc;
        dim = j + 1;
        const double* r = hgm + (long)j * GM_REC;
        double* hc = &H[(size_t)j * (m + 1)];
        for (int i = 0; i <= j; ++i) hc[i] = hs[j] * hs[i] * r[i];
        double nraw = r[GM_REC - 1];
        if (nraw < reorth_fac * std::sqrt(r[j + 1])) {
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
        hc[j + 1] = hs[j] * nraw;
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
      // x += sum_i (y_i sig_i) V_i; the pinned coefficients are not rewritten
      // before the next drain of the stream
      for (int i = 0; i < dim; ++i) hy[i] = y[i] * hs[i];
      ATMO_HIP_CHECK(hipMemcpyAsync(gm_coef.p, hy, sizeof(double) * dim, hipMemcpyHostToDevice, s_));
      gm_update(Ns, x, V, ld, dim, gm_coef.p, s_);
      if (state != 0) break;
    }
    if (state == -1) {
      ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 250, scal.p + 250, sizeof(double), hipMemcpyDeviceToHost, s_));
      sync();
      if (hscal[250] != 0.0) throw_bad_density();
      throw NoConvergence("GMRES: no convergence");
    }
    st_.gmres_its += acc;
    return acc;
  }

  // ---------------------------------------------------- solver stages
  // update_density (Atmospheric_Flow.cc:430-537)
  void update_density(int hs) {
    set_stages(hs, hs == 3 ? 4 : 1);
    double* rhs = gm_p.p;
    if (hs == 1) {
      const double* s[2] = {rho_old.p, u_old.p};
      rhs_rho(s, 2, rhs);
      solve_mass_mg(1, rho2.p, rhs, P.eps);
    } else if (hs == 2) {
      const double* s[4] = {rho_old.p, u_old.p, rho2.p, u2.p};
      rhs_rho(s, 4, rhs);
      solve_mass_mg(1, rho3.p, rhs, P.eps);
    } else {
      const double* s[6] = {rho_old.p, u_old.p, rho2.p, u2.p, rho3.p, u3.p};
      rhs_rho(s, 6, rhs);
      solve_mass_mg(4, rho_curr.p, rhs, P.eps);
    }
  }

  // pressure_fixed_point (Atmospheric_Flow.cc:544-673); rhs_u and
  // tmp_1 = A^{-1} rhs_u do not depend on the fixed-point iterate and are
  // formed in the stage's first iteration only
  void pressure_fixed_point(int hs, bool first) {
    set_stages(hs, 2);
    set_pres_fixed(p_fixed_old.p);
    const long Ns = ns();
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
    if (first) solve_mass_mg(3, tmp1.p, rhs_u0.p, P.a_solve_tol);
    vec_copy(3 * Ns, rhs_u0.p, rhs_u.p, s_);
    vmult_enthalpy(tmp1.p, s_t3.p);
    vec_axpby(Ns, -1.0, s_t3.p, 1.0, rhs_p.p, s_);
    vec_copy(Ns, p_fixed_old.p, p_fixed.p, s_);
    gmres(p_fixed.p, rhs_p.p, P.eps);
  }

  // update_velocity (Atmospheric_Flow.cc:680-781)
  void update_velocity(int hs) {
    const long Ns = ns();
    if (hs == 1 || hs == 2) {
      set_stages(hs, 3);
      vmult_pressure(p_fixed.p, s_t1.p);
      vec_axpby(3 * Ns, -1.0, s_t1.p, 1.0, rhs_u.p, s_);
      solve_mass_mg(3, u_fixed.p, rhs_u.p, P.eps);
    } else {
      set_stages(3, 5);
      const double* su[9] = {rho_old.p, u_old.p, p_old.p, rho2.p, u2.p, p2.p, rho3.p, u3.p, p3.p};
      rhs_velocity(su, 9, rhs_u.p);
      solve_mass_mg(5, u_curr.p, rhs_u.p, P.eps);
    }
  }

  // update_pressure (Atmospheric_Flow.cc:788-871)
  void update_pressure() {
    set_stages(3, 6);
    const double* sp[11] = {rho_old.p, u_old.p, p_old.p, rho2.p, u2.p, p2.p, rho3.p, u3.p, p3.p, rho_curr.p, u_curr.p};
    rhs_pressure(sp, 11, rhs_p.p);
    solve_mass_mg(6, p_old.p, rhs_p.p, P.eps);
  }

  // fixed-point loop with the relative L-infinity stop test on nodal
  // pressure values (Atmospheric_Flow.cc:1135-1153)
  void fixed_point_loop(int hs) {
    const long Ns = ns();
    for (int iter = 0; iter < P.fp_max_iter; ++iter) {
      pressure_fixed_point(hs, iter == 0);
      update_velocity(hs);
      ++st_.fp_its;
      vec_linf(Ns, p_fixed.p, nullptr, scal.p + 2, rb, s_);
      vec_linf(Ns, p_fixed.p, p_fixed_old.p, scal.p + 3, rb, s_);
      ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 2, scal.p + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, s_));
      ATMO_HIP_CHECK(hipMemcpyAsync(hscal + 250, scal.p + 250, sizeof(double), hipMemcpyDeviceToHost, s_));
      sync();
      if (hscal[250] != 0.0) throw_bad_density();
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

}  // namespace

std::unique_ptr<Engine> Engine::create_shell(const ShellDesc& d, const EngineParams& p) {
  return std::unique_ptr<Engine>(new ShellEngine(d, p));
}

size_t shell_device_bytes(const ShellDesc& d, const EngineParams& p) {
  if (d.refinements < 0 || d.refinements > 7 || (d.panels != 6 && d.panels != 12))
    throw std::invalid_argument("shell: refinements in [0, 7] and 6 or 12 panels");
  const size_t n = (size_t)1 << d.refinements, nc = (size_t)d.panels * n * n * n, Ns = nc * 8;
  size_t b = 0;
  b += (size_t)d.panels * 12 * 8 + nc * 6 * 4 + nc * 6;                 // mesh tables
  b += RedBuf::doubles() * 8 + 256 * 8 + (DIAG_PARTIALS + DG_SLOTS) * 8;  // reductions, diagnostics
  b += (nc * 36 * 2 + nc * 8 * 2 + Ns * 3) * 8;                         // blocks, diagonals, rho_fix, pres_fix
  b += (Ns * 3 * 2 + Ns) * 8;                                           // Schur temporaries
  b += ((size_t)(p.gmres_restart + 2) * (Ns + 32) + Ns + 2 * (MAX_GMRES + 2) + (size_t)MAX_GMRES * GM_REC) * 8;
  b += (size_t)RED_MAX_SLOTS * shell_blocks((int)nc) * 8;                // Arnoldi partial dot products
  b += (Ns * 10 + Ns * 3 * 8) * 8;                                      // IMEX state and work vectors
  return b;
}

}  // namespace atmo
