/* C ABI of the MI355X IMEX-DG hot path (libatmo_hip.so).
 *
 * Replaces, on one GPU with device-resident fields, the deal.II-based
 * operator and solver sequence of Atmospheric_Flows:
 *   - EULEROperator (euler_operator.h:24-211): setters, explicit residuals,
 *     B/C actions, vmult per NS_stage (incl. the Schur complement) and
 *     compute_diagonal;
 *   - the EulerSolver stage solves and the IMEX time loop
 *     (Atmospheric_Flow.cc:430-871, 1100-1257).
 *
 * Conventions: every function returns an atmo_status (0 = ATMO_OK); no C++
 * exception crosses the ABI; the library owns device memory; one context
 * owns one HIP stream; calls on one context must be serialized by the
 * caller (like the reference operator, which mutates NS_stage inside
 * apply_add and is not re-entrant, euler_operator.h:91, 2383-2406).
 *
 * Field layout (host and device): [component][cell][dof], cell = j*nx + i
 * (x fastest), dof = jj*(k+1) + ii (x fastest). Velocity has 2 components.
 * Block ids follow the reference's DoF-handler order (Atmospheric_Flow.cc:334-336):
 * 0 = velocity, 1 = pressure, 2 = density.
 */
#ifndef ATMO_HIP_H
#define ATMO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int atmo_status;
#define ATMO_OK 0
#define ATMO_INVALID_ARG 1
#define ATMO_NO_CONVERGENCE 2 /* deal.II SolverControl::NoConvergence */
#define ATMO_NONFINITE 3
#define ATMO_OOM 4
#define ATMO_DEVICE 5

#define ATMO_BLOCK_VELOCITY 0
#define ATMO_BLOCK_PRESSURE 1
#define ATMO_BLOCK_DENSITY 2

typedef struct atmo_ctx atmo_ctx;
typedef struct atmo_vec atmo_vec;

/* Terrain-following structured mesh: vertex (i,j) at x = i*Lx/nx,
 * z = h[i] + (H - h[i]) * j / nz, bilinear cells, walls at bottom and top,
 * lateral boundaries periodic (h[nx] == h[0]) or walls. Replaces the
 * reference's p4est sphere mesh (Atmospheric_Flow.cc:279-289). */
typedef struct {
  int degree; /* FE_DGQ degree k (1..3); quadrature QGauss(2k+1) */
  int nx, nz;
  int periodic;
  double Lx, H;
  const double* h; /* nx+1 terrain heights */
} atmo_mesh_desc;

/* Run-time parameters: the reference's Data_Storage (runtime_parameters.h:16-178)
 * plus its hard-coded solver constants. */
typedef struct {
  double Ma, Fr, dt;           /* Physical data / Time step data */
  double gamma_gas;            /* EquationData::Cp_Cv */
  double grav_x, grav_z;       /* gravity unit vector e_r */
  double grav_factor;          /* radial factor a^2/r^2 (1 for constant gravity) */
  double eps;                  /* Data solve / eps */
  int max_iterations;          /* Data solve / max_iterations */
  double schur_inner_tol;      /* euler_operator.h:2391 (1e-12) */
  int schur_inner_max;         /* euler_operator.h:2391 (10000) */
  double a_solve_tol;          /* Atmospheric_Flow.cc:571 (1e-12) */
  int fp_max_iter;             /* Atmospheric_Flow.cc:1135 (100) */
  double fp_tol;               /* Atmospheric_Flow.cc:1149 (1e-10) */
  int fp_lock;                 /* >0: exactly this many fixed-point iterations */
  int mg_min_cells;            /* coarsest MG level keeps >= this many cells */
  int cheb_degree;             /* Atmospheric_Flow.cc:492 (3) */
  double cheb_range;           /* Atmospheric_Flow.cc:491 (15) */
  int cheb_eig_its;            /* Atmospheric_Flow.cc:493 (10) */
  int gmres_restart;           /* deal.II SolverGMRES default (28) */
  int precision;               /* 0: FP64, the reference's solvers; 1: mixed form: FP32 multigrid inside
                                  FP64 Krylov, FP64 cellwise-direct A^{-1} (inner Schur and A solves);
                                  2: the mixed form's FP64 comparator: the same cellwise-direct A^{-1}
                                  with FP64 multigrid (differs from 1 only in the levels' precision) */
  int mg_matrix_cores;         /* 1: the FP32 multigrid sweeps (precision 1) at degree 2 and 3 run on the
                                  matrix cores (FP32 MFMA sum factorization); 0: on the vector ALUs */
  int arnoldi_waves_per_simd;  /* waves per SIMD the GMRES Arnoldi kernel of the cellwise-direct forms is
                                  compiled for: 3 or 4 (0: 3 at degree 2, else 4); the same results */
} atmo_params;

/* Kernel shapes timed by the autotuners (microseconds per launch) and the choice made. */
typedef struct {
  double arnoldi_us[2];        /* the Arnoldi step (14 basis vectors) at 3 and 4 waves per SIMD */
  double sweep_us[2];          /* the FP32 A-block multigrid sweep on the vector ALUs and on the matrix cores */
  int arnoldi_waves_per_simd;  /* chosen (0: not tuned: AMR mesh, precision 0 or gmres_restart < 13) */
  int mg_matrix_cores;         /* chosen (-1: not tuned: no FP32 multigrid at degree 2 or 3) */
} atmo_kernel_tuning;

typedef struct {
  long cg_its, gmres_its, inner_cg_its, fp_its, mg_cycles, coarse_its, schur_matvecs, host_syncs;
} atmo_stats;

/* Defaults equal to the reference's hard-coded values. */
void atmo_params_default(atmo_params* p);

/* Library version string. */
const char* atmo_version(void);
/* Message of the last failed call on ctx (or of the last failed create when ctx is NULL). */
const char* atmo_last_error(const atmo_ctx* ctx);

/* EulerSolver constructor part: mesh, MG hierarchy, device buffers
 * (Atmospheric_Flow.cc:217-272, setup_dofs :296-409). */
atmo_status atmo_ctx_create(const atmo_mesh_desc* mesh, const atmo_params* params, atmo_ctx** out);
/* Locally refined mesh over the base mesh `base` (refine_global replaced by
 * local quadtree refinement; the reference's triangulation is created with
 * limit_level_difference_at_vertices + construct_multigrid_hierarchy,
 * Atmospheric_Flow.cc:225-226, 286). leaves[5*c .. 5*c+4] = (bi, bj, level,
 * sx, sz) of cell c: base cell (bi, bj), quadtree level, position in
 * [0, 2^level)^2; listed base cell by base cell in the order bj*nx + bi; the
 * cell order of every field. Neighbours may differ by one level at most.
 * Hanging faces are integrated on the smaller side's face (deal.II's
 * subfaces); the multigrid levels coarsen globally down to the base mesh. */
atmo_status atmo_ctx_create_amr(const atmo_mesh_desc* base, const int* leaves, int nleaves, const atmo_params* params,
                                atmo_ctx** out);
atmo_status atmo_ctx_destroy(atmo_ctx* ctx);
/* Hill-flow initial state on the context's own mesh (structured or locally
 * refined), as atmo_init_hill_state. */
atmo_status atmo_ctx_init_hill_state(atmo_ctx* ctx, double theta0, double p0, double N, double U, double g, double R,
                                     double gamma_gas, double perturbation, uint64_t seed, double* rho, double* u,
                                     double* p);
atmo_status atmo_set_params(atmo_ctx* ctx, const atmo_params* params);
atmo_status atmo_ctx_info(atmo_ctx* ctx, int* ncells, long* scalar_dofs, int* nlevels);
atmo_status atmo_sync(atmo_ctx* ctx);
atmo_status atmo_get_stats(atmo_ctx* ctx, atmo_stats* st);
atmo_status atmo_reset_stats(atmo_ctx* ctx);

/* MatrixFree::initialize_dof_vector(v, block) (Atmospheric_Flow.cc:351-376). */
atmo_status atmo_vec_alloc(atmo_ctx* ctx, int block, atmo_vec** out);
atmo_status atmo_vec_free(atmo_ctx* ctx, atmo_vec* v);
atmo_status atmo_vec_size(const atmo_vec* v, size_t* n);
atmo_status atmo_vec_upload(atmo_ctx* ctx, atmo_vec* v, const double* host, size_t n);
atmo_status atmo_vec_download(atmo_ctx* ctx, const atmo_vec* v, double* host, size_t n);
/* Device address of the vector's storage (for zero-copy callers on ctx's stream). */
atmo_status atmo_vec_device_ptr(atmo_vec* v, double** dev);
/* y = a*x + b*y on the device (b == 0 overwrites y): LinearAlgebra::distributed::
 * Vector::equ / add / sadd, e.g. pres_tmp.equ(1, pres_fixed); pres_tmp.add(-1,
 * pres_fixed_old) of the fixed-point test (Atmospheric_Flow.cc:1144-1145).
 * x and y belong to the same block. */
atmo_status atmo_vec_axpby(atmo_ctx* ctx, double a, const atmo_vec* x, double b, atmo_vec* y);
/* Scalar product x.y (Vector::operator*), summed in a fixed order (bitwise
 * reproducible); synchronizes ctx's stream. */
atmo_status atmo_vec_dot(atmo_ctx* ctx, const atmo_vec* x, const atmo_vec* y, double* result);
/* L-infinity norm max_i |x_i|: the fixed-point test's integrate_difference(...,
 * ZeroFunction, quadrature_temperature, Linfty_norm) + compute_global_error
 * (Atmospheric_Flow.cc:1140-1148), whose QGaussLobatto(k+1) points are the
 * FE_DGQ support points, so the point values are the vector's entries; for
 * the velocity block the maximum also runs over both components, as deal.II's
 * Linfty_norm does. Synchronizes ctx's stream. */
atmo_status atmo_vec_linf(atmo_ctx* ctx, const atmo_vec* x, double* result);

/* No counterpart in the reference: a variant of atmo_vec_dot for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The same fixed-order scalar product
 * (bitwise equal to atmo_vec_dot on the same data) on caller-owned device
 * memory, with no context; it allocates nothing: its block partials live in a
 * static device array of the library (1024 doubles, 8 KiB per GPU, held until
 * the process ends). x, y: n doubles on the current HIP device; result: one
 * double of device memory, written by the call's last kernel; stream: a
 * hipStream_t (NULL = the null stream). Returns without synchronizing. Every
 * call on one GPU shares that GPU's array: calls that overlap in time on one
 * GPU race (serialize them, e.g. on one stream). n <= ATMO_DOT_NOALLOC_MAX_N,
 * else ATMO_INVALID_ARG. */
#define ATMO_DOT_NOALLOC_MAX_N (((size_t)1) << 36)
atmo_status atmo_dot_noalloc(const double* x, const double* y, size_t n, void* stream, double* result);

/* EULEROperator::set_HYPERBOLIC_stage / set_NS_stage (euler_operator.h:283-303). */
atmo_status atmo_set_stages(atmo_ctx* ctx, int hyperbolic_stage, int ns_stage);
/* set_rho_for_fixed (:312-315): also refreshes the level densities
 * (interpolate_to_mg, Atmospheric_Flow.cc:1127-1132). */
atmo_status atmo_set_rho_for_fixed(atmo_ctx* ctx, const atmo_vec* rho);
/* set_pres_fixed (:324-327). */
atmo_status atmo_set_pres_fixed(atmo_ctx* ctx, const atmo_vec* p);

/* vmult_rhs_rho_update / vmult_rhs_pressure / vmult_rhs_velocity_fixed
 * (euler_operator.h:697-712, 1364-1381, 2308-2325); src follows the same
 * positional conventions as the reference's callers (Atmospheric_Flow.cc:439-449,
 * 555-568, 696-698, 796-799). */
atmo_status atmo_rhs_rho(atmo_ctx* ctx, atmo_vec* dst, const atmo_vec* const* src, int nsrc);
atmo_status atmo_rhs_pressure(atmo_ctx* ctx, atmo_vec* dst, const atmo_vec* const* src, int nsrc);
atmo_status atmo_rhs_velocity(atmo_ctx* ctx, atmo_vec* dst, const atmo_vec* const* src, int nsrc);
/* vmult_pressure (B, :2420-2435) and vmult_enthalpy (C, :2438-2453). */
atmo_status atmo_vmult_pressure(atmo_ctx* ctx, atmo_vec* dst_u, const atmo_vec* src_p);
atmo_status atmo_vmult_enthalpy(atmo_ctx* ctx, atmo_vec* dst_p, const atmo_vec* src_u);
/* MatrixFreeOperators::Base::vmult = zero + apply_add per NS_stage (:2360-2417);
 * NS_stage 2 is the Schur complement D - C A^{-1} B with its inner CG. */
atmo_status atmo_vmult(atmo_ctx* ctx, atmo_vec* dst, const atmo_vec* src);
/* compute_diagonal + get_matrix_diagonal_inverse (:2589-2626). */
atmo_status atmo_inverse_diagonal(atmo_ctx* ctx, atmo_vec* dst);

/* SolverCG + PreconditionMG on the mass-type block of ns_stage (1/4 M_rho,
 * 3/5 A, 6 D) to tol_factor*|b| (Atmospheric_Flow.cc:452-536, 701-780, 801-870). */
atmo_status atmo_solve_cg_mg(atmo_ctx* ctx, int ns_stage, atmo_vec* x, const atmo_vec* b, double tol_factor,
                             int* iterations);
/* SolverGMRES + PreconditionJacobi on the Schur complement (:656-672). */
atmo_status atmo_solve_schur_gmres(atmo_ctx* ctx, atmo_vec* x, const atmo_vec* b, double tol_factor,
                                   int* iterations);

/* IMEX driver state (rho_old, u_old, pres_old) and time steps (EulerSolver::run body). */
atmo_status atmo_state_upload(atmo_ctx* ctx, const double* rho, const double* u, const double* p);
atmo_status atmo_state_download(atmo_ctx* ctx, double* rho, double* u, double* p);
atmo_status atmo_imex_step(atmo_ctx* ctx, int nsteps);

/* Per-step diagnostics of EulerSolver::run, computed on the device inside
 * every step without host synchronization: get_min_density / get_max_density
 * after each density update (Atmospheric_Flow.cc:1121-1122, 1164-1165,
 * 1205-1206; :1020-1061), compute_max_celerity and get_max_velocity of the new
 * state (:1066-1087, :992-1015) at the QGaussLobatto(k+1) points (the FE_DGQ
 * support points, :86-88, 233-235: the nodal values), and the Courant
 * numbers built from them with GridTools::minimal_cell_diameter (:1225-1235). */
typedef struct {
  double min_density[3];  /* min rho at the points after the density update of stage 1..3 */
  double max_density[3];  /* linfty_norm of the same densities */
  double max_celerity;    /* max sqrt(gamma p / rho); the reference prints it times 1/Ma */
  double max_velocity;    /* max |u| */
  double min_cell_diameter;
  double cfl_c;           /* dt * max_celerity / Ma * k * sqrt(2) / min_cell_diameter */
  double cfl_u;           /* dt * max_velocity * k * sqrt(2) / min_cell_diameter */
  long nonfinite;         /* non-finite values met by these reductions (0 on a healthy state) */
} atmo_step_diag;

/* Reads the record of the last atmo_imex_step (synchronizes the context's
 * stream). of_state != 0 first recomputes the stage-3 density and the flow
 * entries on the current state (e.g. the initial state before any step). */
atmo_status atmo_step_diagnostics(atmo_ctx* ctx, int of_state, atmo_step_diag* out);

/* Hill-flow initial state at the FE_DGQ support points (replaces
 * EquationData's analytic sphere state, equation_data.h:47-176, evaluated by
 * VectorTools::interpolate at Atmospheric_Flow.cc:420-422). */
atmo_status atmo_init_hill_state(const atmo_mesh_desc* mesh, double theta0, double p0, double N, double U, double g,
                                 double R, double gamma_gas, double perturbation, uint64_t seed, double* rho,
                                 double* u, double* p);

/* VTU output of the current state, as EulerSolver::output_results
 * (Atmospheric_Flow.cc:886-918): point data "rho", "u" (3 components, the
 * third zero) and "p" on DataOut's patches of `degree` subdivisions per cell
 * (build_patches(mapping, degree_u), :903), each cell with its own points; with
 * high_order != 0 one Lagrange quadrilateral of (k+1)^2 equispaced nodes per
 * cell instead (the solution_high_order-*.vtu of :909-913). Data arrays are
 * VTK inline binary (base64, UInt32 size header). Off the timed path: the state
 * is copied to the host. */
atmo_status atmo_write_vtu(atmo_ctx* ctx, const char* path, int high_order);
/* The same for a state held on the host, on a structured mesh; needs no GPU. */
atmo_status atmo_write_vtu_host(const atmo_mesh_desc* mesh, const double* rho, const double* u, const double* p,
                                const char* path, int high_order);

/* Measurement helper: launches kernel `which` nrep times back-to-back on the
 * context's stream on its current data and reports the mean duration from
 * HIP events recorded on that stream. which: 0 A apply (k_mass, weighted),
 * 1 fused inner-CG apply (k_jcg_apply), 2 M_rho apply, 3 B, 4 C,
 * 5 stage-3 pressure residual, 6 fused inner-CG update (k_jcg_upd),
 * 7 plain device-to-device copy of one velocity field (calibration),
 * 8 B with the cellwise A^{-1} (k_b_ainv, mixed form only), 9 scaled D - C
 * (k_cellface<PhysDminusC>), 10 GMRES dot products of one vector with 14
 * basis vectors (k_mdot_part + finalize), 11 GMRES projection on 13 basis
 * vectors with the norm of the result (k_maxpy_norm + finalize; needs
 * gmres_restart >= 13 for both), 12 the GMRES Arnoldi step's scaled D - C
 * with the dot products of its result against 14 basis vectors
 * (k_dmc_dots, partial sums only; applied to the newest basis vector, whose
 * staged values serve its own dot product), 13 the same kernel with no basis vector,
 * 14 the FP32 multigrid pre-smoothing sweep with its residual on the finest level
 * (velocity block A, mixed form only) on the matrix cores (k_cheb_sweep_mx), 15 the
 * same sweep on the vector ALUs (k_cheb_sweep).
 * Also reports the kernel's algorithmic HBM bytes per launch. */
atmo_status atmo_time_kernel(atmo_ctx* ctx, int which, int nrep, double* ms_per_launch, double* bytes_per_launch);
/* No counterpart in the reference: the kernel-shape autotuner. Times the tunable
 * kernel shapes on the context's own mesh (nrep launches each, two alternating
 * rounds) — the Arnoldi kernel's waves per SIMD and the FP32 multigrid sweep's
 * form (matrix cores with 16 cells per wave, or vector ALUs) — and sets the
 * fastest in the context's parameters (results do not depend on the choice). */
atmo_status atmo_tune_kernels(atmo_ctx* ctx, int nrep, atmo_kernel_tuning* out);

/* ------------------------------------------------------------ spherical shell
 * The reference's own 3D case (EulerSolver<3>, Atmospheric_Flow.cc:1277):
 * FE_DGQ(1) for density, velocity and pressure with QGauss(3)
 * (Atmospheric_Flow.cc:345) on a shell between two spheres
 * (GridGenerator::concentric_hyper_shells(0.9, 1.0) scaled by a and refined
 * globally, :279-289), MappingQ(2) on the spherical manifold (:239), radial
 * gravity g a^2/r^2 e_r (euler_operator.h:790-811), walls at both radii.
 * A shell context takes every operator, solver and IMEX entry point above;
 * its velocity block has 3 components. Cells: c = ((panel*n + layer)*n +
 * row)*n + column with n = 2^refinements; dofs: d = a + 2b + 4c at the
 * cell vertex (xi, eta, zeta) = (a, b, c), zeta outwards. The mass-type
 * solves (atmo_solve_cg_mg, the inner solve of the Schur complement) are
 * cell-by-cell products with the FP64 block inverses (0 iterations): the
 * exact solutions the reference's CG + MG approach; every precision
 * setting runs this FP64 form. Replaces the mesh of
 * Atmospheric_Flow.cc:279-289 and the solver set-up of :296-409 for dim = 3. */
typedef struct {
  int refinements; /* Space discretization / n_of_refines (parameter-file.prm:19), 0..7 */
  int panels;      /* coarse cells per shell: 6 (cube, hyper_shell's default) or 12 (rhombic dodecahedron) */
  double r_inner, r_outer; /* shell radii: 0.9 a and a in the reference (Atmospheric_Flow.cc:283-284) */
  double planet;   /* a of the gravity factor a^2/r^2 (EquationData::a, equation_data.h:27) */
} atmo_shell_desc;

atmo_status atmo_ctx_create_shell(const atmo_shell_desc* shell, const atmo_params* params, atmo_ctx** out);
/* Velocity components of the context's fields: 2 (terrain), 3 (shell). */
atmo_status atmo_ctx_dim(atmo_ctx* ctx, int* dim);
/* Physical positions of the dofs of a shell context, [cell][8][3]. */
atmo_status atmo_shell_dof_points(atmo_ctx* ctx, double* xyz);
/* The solid-body-rotation steady state (EquationData::Density/Velocity/
 * Pressure, equation_data.h:47-176) at the dofs, in the reference's
 * non-dimensional variables (rho/rho_ref = p/p_ref = exp(q), u = Omega
 * (-y, x, 0)); RT = R T_bar. VectorTools::interpolate, Atmospheric_Flow.cc:420-422. */
atmo_status atmo_shell_sbr_state(atmo_ctx* ctx, double Omega, double a, double g, double RT, double* rho, double* u,
                                 double* p);
/* analyze_results (Atmospheric_Flow.cc:926-985) on the context's state: L2
 * errors against the steady state with QGauss(2) on the mapped cells;
 * out[6] = {u, u relative, rho, rho relative, p, p relative}. Copies the
 * state to the host. */
atmo_status atmo_shell_l2_errors(atmo_ctx* ctx, double Omega, double a, double g, double RT, double* out);
/* Device bytes a shell context of this description and parameters allocates. */
atmo_status atmo_shell_device_bytes(const atmo_shell_desc* shell, const atmo_params* params, size_t* bytes);

/* ------------------------------------------------------------ device footprint
 * No counterpart in the reference (it sizes nothing up front): device bytes
 * a context of this mesh and parameters allocates, for admission decisions
 * before creating it; and what an existing context allocated. */
atmo_status atmo_device_bytes(const atmo_mesh_desc* mesh, const atmo_params* params, size_t* bytes);
atmo_status atmo_ctx_device_bytes(atmo_ctx* ctx, size_t* bytes);

/* ------------------------------------------------------------ ensemble / serving
 * No counterpart in the reference, which runs one simulation per process:
 * many independent members (own mesh, parameters, hill initial state, step
 * count, arrival time) run concurrently on one GPU. `slots` host threads
 * each drive one member's context on its own HIP stream; requests are
 * admitted in arrival order when a slot is free and their device footprint
 * (atmo_device_bytes) fits the remaining HBM budget; an arrived request that
 * does not fit is passed by later ones that do. Step and request latencies
 * are measured on the host clock. A member's results do not depend on what
 * else runs: every kernel and reduction is deterministic and per-stream. */
typedef struct atmo_ensemble atmo_ensemble;

typedef struct {
  atmo_mesh_desc mesh;    /* h is copied at submit */
  atmo_params params;
  double theta0, p0, N, U, g, R, gamma_gas, perturbation; /* hill initial state (atmo_init_hill_state) */
  uint64_t seed;
  int nsteps;
  double arrival_s;       /* seconds after atmo_ensemble_run starts */
} atmo_member_desc;

typedef struct {
  int status;             /* atmo_status of the member's run */
  double arrival_s, admit_s, start_s, finish_s; /* seconds after the run started */
  int steps_done;
  size_t device_bytes;
  atmo_stats stats;
  char error[256];
} atmo_member_result;

atmo_status atmo_ensemble_create(int slots, double hbm_budget_bytes, int keep_states, atmo_ensemble** out);
atmo_status atmo_ensemble_destroy(atmo_ensemble* e);
atmo_status atmo_ensemble_submit(atmo_ensemble* e, const atmo_member_desc* m, int* id);
/* No counterpart in the reference: the batch-shape autotuner. Runs the first
 * sample_members submitted members (arriving at once) for probe_steps steps
 * with each of the ncand candidate slot counts (1..64), measures the
 * aggregate DoF-steps/s of the stepping phase, and keeps as the ensemble's
 * slot count (*best_slots) the fewest slots whose rate is within
 * rate_tolerance (a fraction in [0, 1); 0 = the fastest candidate) of the
 * best rate: fewer concurrent members step with lower latency. rates (may be
 * NULL) receives each candidate's rate (0 when a member of that probe failed). */
atmo_status atmo_ensemble_autotune(atmo_ensemble* e, const int* candidates, int ncand, int probe_steps,
                                   int sample_members, double rate_tolerance, int* best_slots, double* rates);
/* No counterpart in the reference: the kernel-shape autotuner of an ensemble. For
 * every class of submitted members with one degree, mesh size and precision, times
 * the tunable kernel shapes on one engine of the class (as atmo_tune_kernels) and
 * sets the choices for every member of the class; out[c] (c < cap) receives class
 * c's timings and choices, classes in order of first appearance, and *nclasses
 * their number. */
atmo_status atmo_ensemble_tune_kernels(atmo_ensemble* e, int nrep, atmo_kernel_tuning* out, int cap, int* nclasses);
/* Runs every submitted member; returns when all are done (wall_s may be NULL). */
atmo_status atmo_ensemble_run(atmo_ensemble* e, double* wall_s);
/* Result of member id; step_ms (may be NULL) receives up to cap per-step wall times. */
atmo_status atmo_ensemble_result(atmo_ensemble* e, int id, atmo_member_result* out, double* step_ms, int cap);
/* Final state of member id (keep_states != 0); ns = scalar dofs per field. */
atmo_status atmo_ensemble_state(atmo_ensemble* e, int id, double* rho, double* u, double* p, size_t ns);

#ifdef __cplusplus
}
#endif

#endif /* ATMO_HIP_H */
