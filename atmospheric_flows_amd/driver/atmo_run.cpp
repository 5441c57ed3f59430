// Command-line driver: the reference's main + EulerSolver::run
// (Atmospheric_Flow.cc:1100-1257, 1264-1311) on the terrain-following hill
// case, through the C++ operator interface of libatmo_hip.so.
//
//   atmo_run [parameter-file.prm]          run from initial_time to final_time
//   atmo_run --print-params <file.prm>     parse only, print every field
//
// Per step it prints what the reference prints (step/time, min/max density
// after each density update, maximum celerity and velocity with the
// acoustic and advective Courant numbers), clips the last time step to land
// on final_time (:1243-1250) and stops with an error on a non-finite state.
// Field output as output_results (:886-918, 1104, 1237-1256): <dir>/solution-NNNNN.vtu
// and <dir>/solution_high_order-NNNNN.vtu at step 0, every output_interval steps
// and after the last step; written only when <dir> exists (the reference's
// output directory must exist beforehand as well, runtime_parameters.h:46-48).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/atmo_operator.hpp"
#include "prm.hpp"

namespace {

int run(const atmo::DataStorage& d) {
  std::vector<double> h(d.nx + 1);
  for (int i = 0; i <= d.nx; ++i) {
    const double x = d.Lx * i / d.nx, s = (x - d.hill_center) / d.hill_half_width;
    h[i] = d.hill_height / (1.0 + s * s);
  }
  atmo_mesh_desc mesh{d.degree, d.nx, d.nz, 1, d.Lx, d.H, h.data()};
  atmo_params p = atmo::default_params();
  p.Ma = d.Mach;
  p.Fr = d.Froude;
  p.dt = d.dt;
  p.eps = d.eps;
  p.max_iterations = d.max_iterations;
  p.precision = d.precision == "mixed" ? 1 : (d.precision == "fp64_direct" ? 2 : 0);
  atmo::Context ctx(mesh, p);
  std::printf("Number of cells %d, scalar DoFs per field %ld, MG levels %d, precision %s\n", ctx.n_cells(),
              ctx.scalar_dofs(), ctx.n_levels(), d.precision.c_str());

  std::vector<double> rho, u, pr;
  ctx.initial_state(rho, u, pr, 300.0, 1e5, d.brunt_vaisala, d.wind, d.gravity, 287.0, 1.4, 1e-4,
                    (uint64_t)d.seed);
  ctx.upload_state(rho, u, pr);
  struct stat sb;
  const bool out = stat(d.dir.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode);
  if (!out) std::printf("Output directory %s does not exist: no VTU output\n", d.dir.c_str());
  auto output_results = [&](unsigned int step) {
    if (!out) return;
    char num[16];
    std::snprintf(num, sizeof num, "%05u", step);
    ctx.write_vtu(d.dir + "/solution-" + num + ".vtu", false);
    ctx.write_vtu(d.dir + "/solution_high_order-" + num + ".vtu", true);
  };
  output_results(0);

  const double T = d.final_time;
  double time = d.initial_time, dt = d.dt;
  unsigned int n = 0;
  double wall = 0.0;
  while (std::abs(T - time) > 1e-10) {
    time += dt;
    ++n;
    std::printf("Step = %u Time = %.10g\n", n, time);
    const auto t0 = std::chrono::steady_clock::now();
    ctx.imex_step(1);
    wall += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    atmo_step_diag g{};
    atmo::check(atmo_step_diagnostics(ctx.handle(), 0, &g), ctx.handle());
    if (d.verbose)
      for (int s = 0; s < 3; ++s)
        std::printf("  stage %d: Minimum density %.10g Maximum density %.10g\n", s + 1, g.min_density[s],
                    g.max_density[s]);
    std::printf("Maximum celerity = %.10g\nCFL_c = %.10g\nMaximum velocity = %.10g\nCFL_u = %.10g\n",
                g.max_celerity / d.Mach, g.cfl_c, g.max_velocity, g.cfl_u);
    if (g.nonfinite != 0) throw atmo::Error(ATMO_NONFINITE, "non-finite values in the state after step " +
                                                                std::to_string(n));
    if (n % d.output_interval == 0) output_results(n);
    if (T - time < dt && T - time > 1e-10) {
      dt = T - time;
      atmo_params q = ctx.params();
      q.dt = dt;
      ctx.set_params(q);
    }
  }
  if (n % d.output_interval != 0) output_results(n);
  const atmo_stats st = ctx.stats();
  const double dofs = 4.0 * (double)ctx.scalar_dofs();
  std::printf("%u steps, %.3f s in steps, %.4g DoF-steps/s; CG %ld GMRES %ld inner CG %ld fixed-point %ld\n", n,
              wall, wall > 0 ? dofs * n / wall : 0.0, st.cg_its, st.gmres_its, st.inner_cg_its, st.fp_its);
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  try {
    if (argc == 3 && std::string(argv[1]) == "--print-params") {
      atmo::DataStorage d;
      d.read_data(argv[2]);
      std::cout << d.dump();
      return 0;
    }
    if (argc > 2) {
      std::cerr << "usage: atmo_run [parameter-file.prm] | --print-params <file.prm>\n";
      return 2;
    }
    atmo::DataStorage d;
    d.read_data(argc == 2 ? argv[1] : "parameter-file.prm");
    const int rc = run(d);
    std::cout << "----------------------------------------------------\n"
              << "Apparently everything went fine!\n";
    return rc;
  } catch (std::exception& exc) {
    std::cerr << "\n----------------------------------------------------\n"
              << "Exception on processing: \n"
              << exc.what() << "\nAborting!\n"
              << "----------------------------------------------------\n";
    return 1;
  }
}
