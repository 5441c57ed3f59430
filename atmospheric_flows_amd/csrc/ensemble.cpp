// Ensemble / serving loop (see ensemble.h).
#include "ensemble.h"

#include <algorithm>
#include <chrono>
#include <new>
#include <stdexcept>
#include <thread>

#include "../../include/atmo_hip.h"

namespace atmo {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int status_of(const std::exception& e) {
  if (dynamic_cast<const std::invalid_argument*>(&e)) return ATMO_INVALID_ARG;
  if (dynamic_cast<const std::bad_alloc*>(&e)) return ATMO_OOM;
  const std::string w = e.what();
  if (w.find("out of memory") != std::string::npos) return ATMO_OOM;
  if (w.find("no convergence") != std::string::npos) return ATMO_NO_CONVERGENCE;
  if (w.find("non-finite") != std::string::npos) return ATMO_NONFINITE;
  return ATMO_DEVICE;
}

}  // namespace

Ensemble::Ensemble(int slots, double hbm_budget_bytes, bool keep_states)
    : slots_(slots), budget_(hbm_budget_bytes), keep_(keep_states) {
  if (slots < 1 || slots > 64) throw std::invalid_argument("ensemble: slots must be 1..64");
  if (!(hbm_budget_bytes > 0)) throw std::invalid_argument("ensemble: HBM budget must be positive");
}

int Ensemble::submit(const MemberSpec& m) {
  if (m.nsteps < 0) throw std::invalid_argument("ensemble: nsteps must be >= 0");
  if (m.arrival_s < 0) throw std::invalid_argument("ensemble: arrival time must be >= 0");
  if ((int)m.mesh.h.size() != m.mesh.nx + 1) throw std::invalid_argument("ensemble: need nx+1 terrain heights");
  need_.push_back(engine_device_bytes(m.degree, m.mesh, m.params));  // validates degree and sizes
  specs_.push_back(m);
  res_.emplace_back();
  res_.back().arrival_s = m.arrival_s;
  return (int)specs_.size() - 1;
}

const MemberResult& Ensemble::result(int id) const {
  if (id < 0 || id >= (int)res_.size()) throw std::invalid_argument("ensemble: unknown member id");
  return res_[id];
}

void Ensemble::run_member(int id) {
  const MemberSpec& s = specs_[id];
  MemberResult& r = res_[id];
  try {
    auto eng = Engine::create(s.degree, s.mesh, s.params);
    r.device_bytes = eng->device_bytes();
    const long Ns = eng->ns();
    std::vector<double> rho(Ns), u(2 * Ns), p(Ns);
    hill_initial_state(s.degree, s.mesh, s.theta0, s.p0, s.Nbv, s.U, s.g, s.Rgas, s.gamma_gas, s.pert, s.seed,
                       rho.data(), u.data(), p.data());
    hipStream_t st = eng->stream();
    ATMO_HIP_CHECK(hipMemcpyAsync(eng->state_ptr(0), rho.data(), Ns * sizeof(double), hipMemcpyHostToDevice, st));
    ATMO_HIP_CHECK(hipMemcpyAsync(eng->state_ptr(1), u.data(), 2 * Ns * sizeof(double), hipMemcpyHostToDevice, st));
    ATMO_HIP_CHECK(hipMemcpyAsync(eng->state_ptr(2), p.data(), Ns * sizeof(double), hipMemcpyHostToDevice, st));
    ATMO_HIP_CHECK(hipStreamSynchronize(st));
    r.start_s = now_s() - t0_;
    r.step_ms.reserve(s.nsteps);
    for (int n = 0; n < s.nsteps; ++n) {
      const double t = now_s();
      eng->step();
      ATMO_HIP_CHECK(hipStreamSynchronize(st));
      r.step_ms.push_back(1e3 * (now_s() - t));
    }
    r.stats = eng->stats();
    if (keep_) {
      r.rho.resize(Ns);
      r.u.resize(2 * Ns);
      r.p.resize(Ns);
      ATMO_HIP_CHECK(hipMemcpyAsync(r.rho.data(), eng->state_ptr(0), Ns * sizeof(double), hipMemcpyDeviceToHost, st));
      ATMO_HIP_CHECK(hipMemcpyAsync(r.u.data(), eng->state_ptr(1), 2 * Ns * sizeof(double), hipMemcpyDeviceToHost, st));
      ATMO_HIP_CHECK(hipMemcpyAsync(r.p.data(), eng->state_ptr(2), Ns * sizeof(double), hipMemcpyDeviceToHost, st));
      ATMO_HIP_CHECK(hipStreamSynchronize(st));
    }
  } catch (const std::exception& e) {
    r.status = status_of(e);
    r.error = e.what();
  } catch (...) {
    r.status = ATMO_DEVICE;
    r.error = "unknown error";
  }
  r.finish_s = now_s() - t0_;
}

void Ensemble::worker() {
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    if (pending_.empty()) return;
    const double now = now_s() - t0_;
    int pick = -1;
    double next_arrival = -1;
    for (size_t i = 0; i < pending_.size(); ++i) {
      const int id = pending_[i];
      if (specs_[id].arrival_s > now) {
        next_arrival = specs_[id].arrival_s;
        break;  // pending_ is in arrival order
      }
      if ((double)need_[id] <= budget_ - in_use_) {
        pick = (int)i;
        break;
      }
    }
    if (pick < 0) {
      if (next_arrival >= 0)
        cv_.wait_for(lk, std::chrono::duration<double>(next_arrival - now));
      else
        cv_.wait(lk);  // every arrived request waits for HBM: a finishing member frees some
      continue;
    }
    const int id = pending_[pick];
    pending_.erase(pending_.begin() + pick);
    in_use_ += (double)need_[id];
    ++running_;
    res_[id].admit_s = now;
    lk.unlock();
    run_member(id);
    lk.lock();
    in_use_ -= (double)need_[id];
    --running_;
    res_[id].done = true;
    cv_.notify_all();
  }
}

void Ensemble::run() {
  pending_.clear();
  for (int id = 0; id < (int)specs_.size(); ++id) {
    if (res_[id].done) continue;
    if ((double)need_[id] > budget_) {  // can never be admitted
      res_[id].status = ATMO_OOM;
      res_[id].error = "member needs more device memory than the ensemble's HBM budget";
      res_[id].done = true;
      continue;
    }
    pending_.push_back(id);
  }
  std::stable_sort(pending_.begin(), pending_.end(),
                   [&](int a, int b) { return specs_[a].arrival_s < specs_[b].arrival_s; });
  t0_ = now_s();
  in_use_ = 0;
  running_ = 0;
  std::vector<std::thread> th;
  const int n = std::min<int>(slots_, (int)pending_.size());
  for (int i = 0; i < n; ++i) th.emplace_back([this]() { worker(); });
  for (auto& t : th) t.join();
  wall_s_ = now_s() - t0_;
}

int Ensemble::autotune(const std::vector<int>& candidates, int probe_steps, int sample, double rate_tolerance,
                       std::vector<double>& rates) {
  if (candidates.empty()) throw std::invalid_argument("autotune: no candidates");
  if (probe_steps < 1 || sample < 1) throw std::invalid_argument("autotune: probe_steps and sample must be >= 1");
  if (!(rate_tolerance >= 0.0 && rate_tolerance < 1.0))
    throw std::invalid_argument("autotune: rate_tolerance must be in [0, 1)");
  for (int c : candidates)
    if (c < 1 || c > 64) throw std::invalid_argument("autotune: candidate slot counts must be 1..64");
  const int n = std::min<int>(sample, (int)specs_.size());
  if (n < 1) throw std::invalid_argument("autotune: submit members first");
  rates.assign(candidates.size(), 0.0);
  double best_rate = 0;
  for (size_t ci = 0; ci < candidates.size(); ++ci) {
    Ensemble probe(candidates[ci], budget_, false);
    double dof_steps = 0;
    for (int i = 0; i < n; ++i) {
      MemberSpec m = specs_[i];
      m.nsteps = probe_steps;
      m.arrival_s = 0.0;
      probe.submit(m);
      const double nd = m.degree + 1;
      dof_steps += 4.0 * m.mesh.nx * m.mesh.nz * nd * nd * probe_steps;
    }
    probe.run();
    double t_first = 1e300, t_last = 0;
    bool ok = true;
    for (int i = 0; i < n; ++i) {
      const MemberResult& r = probe.result(i);
      ok = ok && r.status == 0;
      t_first = std::min(t_first, r.start_s);
      t_last = std::max(t_last, r.finish_s);
    }
    rates[ci] = ok && t_last > t_first ? dof_steps / (t_last - t_first) : 0.0;
    best_rate = std::max(best_rate, rates[ci]);
  }
  // the fewest slots within the tolerance of the best rate (the slot count unchanged if every probe failed)
  int best = slots_;
  if (best_rate > 0) {
    best = 0;
    for (size_t ci = 0; ci < candidates.size(); ++ci)
      if (rates[ci] > 0 && rates[ci] >= (1.0 - rate_tolerance) * best_rate && (best == 0 || candidates[ci] < best))
        best = candidates[ci];
  }
  slots_ = best;
  return best;
}

std::vector<Ensemble::ClassTuning> Ensemble::tune_kernels(int nrep) {
  if (nrep < 1) throw std::invalid_argument("tune_kernels: nrep must be >= 1");
  std::vector<ClassTuning> out;
  std::vector<int> cls(specs_.size(), -1);
  for (size_t i = 0; i < specs_.size(); ++i) {
    const MemberSpec& m = specs_[i];
    for (size_t c = 0; c < out.size(); ++c)
      if (out[c].degree == m.degree && out[c].nx == m.mesh.nx && out[c].nz == m.mesh.nz &&
          out[c].precision == m.params.precision)
        cls[i] = (int)c;
    if (cls[i] < 0) {
      cls[i] = (int)out.size();
      out.push_back(ClassTuning{m.degree, m.mesh.nx, m.mesh.nz, m.params.precision, 0, KernelTuning()});
      auto eng = Engine::create(m.degree, m.mesh, m.params);
      out.back().kt = eng->tune_kernels(nrep);
    }
    ++out[cls[i]].members;
  }
  for (size_t i = 0; i < specs_.size(); ++i) {
    const KernelTuning& kt = out[cls[i]].kt;
    if (kt.arnoldi_waves) specs_[i].params.arnoldi_waves = kt.arnoldi_waves;
    if (kt.mg_matrix_cores >= 0) specs_[i].params.mg_matrix_cores = kt.mg_matrix_cores;
  }
  return out;
}

}  // namespace atmo
