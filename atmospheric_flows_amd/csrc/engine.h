// Device engine: the EULEROperator state (euler_operator.h:24-211) and the
// EulerSolver stage sequence (Atmospheric_Flow.cc:430-871, 1100-1257) with
// all fields resident in HBM.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "ops.h"
#include "shell.h"

namespace atmo {

struct EngineParams {
  double Ma = 1.0, Fr = 1.0, dt = 1.0;
  double gamma_gas = 1.4;          // EquationData::Cp_Cv (equation_data.h:24)
  double grav[2] = {0.0, 1.0};     // gravity unit vector (constant for the terrain case)
  double grav_factor = 1.0;        // radial factor a^2/r^2 (1 for constant gravity)
  double eps = 1e-10;              // Data solve / eps (parameter-file.prm:25)
  int max_its = 30000;             // Data solve / max_iterations (:24)
  double schur_inner_tol = 1e-12;  // euler_operator.h:2391
  int schur_inner_max = 10000;     // euler_operator.h:2391
  double a_solve_tol = 1e-12;      // Atmospheric_Flow.cc:571
  int fp_max_iter = 100;           // Atmospheric_Flow.cc:1135
  double fp_tol = 1e-10;           // Atmospheric_Flow.cc:1149
  int fp_lock = 0;                 // >0: run exactly this many fixed-point iterations
  int mg_min_cells = 8;            // coarsest level keeps >= this many cells
  int cheb_degree = 3;             // Atmospheric_Flow.cc:492
  double cheb_range = 15.0;        // :491
  int cheb_eig_its = 10;           // :493
  int gmres_restart = 28;          // deal.II SolverGMRES default (30 temporary vectors)
  // 0: FP64 with the reference's solvers; 1: mixed (FP32 MG levels, cellwise-direct FP64 A^{-1});
  // 2: the mixed form's FP64 comparator (FP64 MG levels, the same cellwise-direct A^{-1})
  int precision = 0;
  // FP32 multigrid sweeps at k = 2, 3 on the matrix cores (MFMA, 16 cells per wave) instead of the vector
  // ALUs (9 or 12 cells per wave)
  int mg_matrix_cores = 1;
  // waves per SIMD the Arnoldi kernel is compiled for (3 or 4; 0: the default for the degree)
  int arnoldi_waves = 0;
};

// Kernel shapes the tuner times on a context's own mesh (Engine::tune_kernels):
// microseconds per launch of each candidate and the choice made.
struct KernelTuning {
  double arnoldi_us[2] = {0, 0};  // the Arnoldi step (14 basis vectors) at 3 and 4 waves per SIMD
  double sweep_us[2] = {0, 0};    // the FP32 A-block sweep on the vector ALUs and on the matrix cores
  int arnoldi_waves = 0;          // chosen (0: not tuned, the form does not run the kernel)
  int mg_matrix_cores = -1;       // chosen (-1: not tuned, no FP32 multigrid at k = 2, 3)
};

// One cell of a locally refined mesh: base cell (bi, bj), quadtree level l,
// position (sx, sz) in [0, 2^l)^2 inside the base cell (amr.h).
struct QCell {
  int bi, bj, l, sx, sz;
};

struct MeshDesc {
  int nx = 0, nz = 0, periodic = 1;
  double Lx = 0, H = 0;
  std::vector<double> h;  // terrain height at the nx+1 vertex columns
};

struct EngineStats {
  long cg_its = 0, gmres_its = 0, inner_cg_its = 0, fp_its = 0, mg_cycles = 0, coarse_its = 0, schur_matvecs = 0;
  long host_syncs = 0;
};

// Block selected by the reference's NS_stage: 1/4 -> M_rho, 3/5 -> A, 2/6 -> D.
enum BlockKind { BLK_RHO = 0, BLK_A = 1, BLK_D = 2 };

class Engine {
 public:
  virtual ~Engine() = default;
  static std::unique_ptr<Engine> create(int k, const MeshDesc& mesh, const EngineParams& p);
  // locally refined mesh over a base mesh; leaves as for build_amr_levels (amr.h)
  static std::unique_ptr<Engine> create_amr(int k, const MeshDesc& base, const std::vector<QCell>& leaves,
                                            const EngineParams& p);
  // the reference's 3D case: FE_DGQ(1) on the spherical shell (shell.h)
  static std::unique_ptr<Engine> create_shell(const ShellDesc& d, const EngineParams& p);

  virtual int dim() const { return 2; }  // velocity components
  virtual const ShellMesh* shell() const { return nullptr; }

  virtual int degree() const = 0;
  virtual int ncells() const = 0;
  virtual long ns() const = 0;  // scalar dofs per field
  virtual int nlevels() const = 0;
  virtual hipStream_t stream() const = 0;
  virtual void set_params(const EngineParams& p) = 0;
  virtual const EngineParams& params() const = 0;
  virtual EngineStats& stats() = 0;
  virtual size_t device_bytes() const = 0;  // device memory requested by this engine
  // Finest-level cell geometry (6 per cell: x0, dx, z00, z10, z01, z11) of a
  // locally refined mesh, empty for a structured one; its minimal diameter.
  virtual const std::vector<double>& cell_geometry() const = 0;
  virtual double min_cell_diameter() const = 0;

  // ---- EULEROperator interface on device pointers
  virtual void set_stages(int hyperbolic_stage, int ns_stage) = 0;
  virtual void set_rho_for_fixed(const double* rho) = 0;  // copies, refreshes level densities
  virtual void set_pres_fixed(const double* p) = 0;       // copies
  virtual void rhs_rho(const double* const* src, int nsrc, double* dst) = 0;
  virtual void rhs_pressure(const double* const* src, int nsrc, double* dst) = 0;
  virtual void rhs_velocity(const double* const* src, int nsrc, double* dst) = 0;
  virtual void vmult_pressure(const double* p, double* dst_u) = 0;   // B
  virtual void vmult_enthalpy(const double* u, double* dst_p) = 0;   // C
  virtual void vmult(const double* src, double* dst) = 0;            // per NS_stage (2: Schur)
  virtual void inverse_diagonal(double* dst) = 0;                    // per NS_stage
  virtual int solve_mass_mg(int ns, double* x, const double* b, double tol_factor) = 0;
  virtual int solve_schur(double* x, const double* b, double tol_factor) = 0;

  // ---- vector operations on n doubles of device memory, on the engine's stream
  virtual void axpby(long n, double a, const double* x, double b, double* y) = 0;  // y = a x + b y
  virtual double dot(long n, const double* x, const double* y) = 0;               // synchronizes
  virtual double linf(long n, const double* x) = 0;                                // max |x_i|; synchronizes

  // ---- IMEX driver (state is the engine's own)
  virtual void load_state(const double* rho, const double* u, const double* p) = 0;  // device pointers
  virtual void store_state(double* rho, double* u, double* p) = 0;
  virtual void step() = 0;
  virtual double* state_ptr(int which) = 0;  // 0 rho, 1 u, 2 p (device)
  // Diagnostics record (DiagSlot layout, DG_SLOTS doubles) of the last step,
  // or, with of_state, first recomputed on the current state (density
  // slots of stage 3 and the flow slots). Synchronizes.
  virtual void step_diagnostics(double* rec, bool of_state) = 0;

  // ---- measurement: mean duration of one kernel on the engine's data
  virtual void time_kernel(int which, int nrep, double* ms_per_launch, double* bytes_per_launch) = 0;
  // ---- autotuner: times the tunable kernel shapes on this engine's mesh and sets the fastest (the
  // results do not depend on the choice); the terrain engines of the cellwise-direct forms
  virtual KernelTuning tune_kernels(int nrep) { return KernelTuning(); }
};

// Device bytes an engine of this mesh and parameters requests (equals
// device_bytes() of such an engine; used for admission before creating it).
size_t engine_device_bytes(int k, const MeshDesc& mesh, const EngineParams& p);
size_t shell_device_bytes(const ShellDesc& d, const EngineParams& p);

// Host-side helpers shared by the C ABI and the driver.
void hill_initial_state(int k, const MeshDesc& m, double theta0, double p0, double Nbv, double U, double g,
                        double Rgas, double gamma_gas, double pert, uint64_t seed, double* rho, double* u,
                        double* p);
// the same on cells given by their geometry (6 per cell, Engine::cell_geometry)
void hill_initial_state_cells(int k, const std::vector<double>& cgeo, double theta0, double p0, double Nbv, double U,
                              double g, double Rgas, double gamma_gas, double pert, uint64_t seed, double* rho,
                              double* u, double* p);

}  // namespace atmo
