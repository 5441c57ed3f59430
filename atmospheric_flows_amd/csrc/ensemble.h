// Ensemble / serving form of the IMEX step: many independent members
// (each its own mesh, parameters, initial state and number of steps) run
// concurrently on one GPU. Requests carry an arrival time; a fixed pool of
// slots (one host thread + one engine on its own HIP stream per running
// member) admits them in arrival order as soon as a slot is free and their
// device footprint fits the HBM budget, skipping past a waiting request that
// does not fit (continuous batching). Per-step and per-request latencies are
// recorded on the host clock.
#pragma once
#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"

namespace atmo {

struct MemberSpec {
  int degree = 3;
  MeshDesc mesh;
  EngineParams params;
  // hill-flow initial state
  double theta0 = 300.0, p0 = 1e5, Nbv = 0.01, U = 10.0, g = 9.81, Rgas = 287.0, gamma_gas = 1.4, pert = 1e-4;
  uint64_t seed = 1234;
  int nsteps = 1;
  double arrival_s = 0.0;  // seconds after the start of run()
};

struct MemberResult {
  int status = 0;  // 0 ok, otherwise an atmo status code
  std::string error;
  double arrival_s = 0, admit_s = 0, start_s = 0, finish_s = 0;  // seconds after the start of run()
  std::vector<double> step_ms;                                    // wall time of each step
  size_t device_bytes = 0;
  EngineStats stats;
  std::vector<double> rho, u, p;  // final state (kept when requested)
  bool done = false;
};

class Ensemble {
 public:
  Ensemble(int slots, double hbm_budget_bytes, bool keep_states);
  int submit(const MemberSpec& m);  // returns the member id
  void run();                       // runs every submitted member, returns when all are done
  const MemberResult& result(int id) const;
  int size() const { return (int)specs_.size(); }
  double wall_s() const { return wall_s_; }
  int slots() const { return slots_; }
  // Batch-shape autotuner: runs the first `sample` submitted members (all
  // arriving at once) for `probe_steps` steps with each candidate number of
  // slots, measures the aggregate DoF-steps/s of the stepping phase (from the
  // first member's first step to the last member's last), keeps as this
  // ensemble's slot count the fewest slots whose rate is within
  // rate_tolerance (a fraction, 0 <= t < 1) of the best rate (0: the fastest
  // candidate; fewer concurrent members step with lower latency) and returns
  // it; rates[i] is candidate i's rate (0 when a member of the probe failed).
  int autotune(const std::vector<int>& candidates, int probe_steps, int sample, double rate_tolerance,
               std::vector<double>& rates);
  // Kernel-shape autotuner: for every class of submitted members with the same degree, mesh size and
  // precision, creates one engine of the class (the first member's mesh and parameters), times the
  // tunable kernel shapes on it (Engine::tune_kernels) and sets the choices in the parameters of every
  // member of the class. Returns one record per class, in order of first appearance.
  struct ClassTuning {
    int degree, nx, nz, precision, members;
    KernelTuning kt;
  };
  std::vector<ClassTuning> tune_kernels(int nrep);

 private:
  void worker();
  void run_member(int id);

  int slots_;
  double budget_;
  bool keep_;
  std::vector<MemberSpec> specs_;
  std::vector<MemberResult> res_;
  std::vector<size_t> need_;
  // scheduler state
  std::mutex mu_;
  std::condition_variable cv_;
  std::vector<int> pending_;  // ids in arrival order, not yet admitted
  double in_use_ = 0;
  int running_ = 0;
  double t0_ = 0, wall_s_ = 0;
};

}  // namespace atmo
