// Reader of the reference's parameter file (runtime_parameters.h:16-178):
// the same subsections, keys, defaults and pattern checks as Data_Storage
// in deal.II ParameterHandler syntax (`subsection`, `set key = value`,
// `end`, `#` comments). Unknown keys, malformed lines and out-of-pattern
// values throw std::runtime_error naming the file and line. A `Device`
// subsection carries the settings of the terrain-following device run.
#pragma once
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace atmo {

struct DataStorage {
  // Physical data / Time step data / Space discretization / Data solve / top level
  double initial_time = 0.0, final_time = 1.0, Mach = 1.0, Froude = 1.0, dt = 5e-4;
  int n_global_refines = 3, max_iterations = 1000;
  double eps = 1e-12;
  bool verbose = true;
  int output_interval = 1;
  std::string dir = "SimTest";
  // Device
  int degree = 3, nx = 512, nz = 256;
  double Lx = 40000.0, H = 20000.0, hill_height = 400.0, hill_half_width = 1000.0, hill_center = 20000.0;
  std::string precision = "fp64";
  double gravity = 9.81, wind = 10.0, brunt_vaisala = 0.01;
  long seed = 20240601;

  // Data_Storage::read_data (runtime_parameters.h:137-176)
  void read_data(const std::string& filename);
  // "key = value" lines of every field, in declaration order
  std::string dump() const;
};

namespace prm_detail {

inline std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

inline bool full_parse_double(const std::string& t, double& v) {
  const char* c = t.c_str();
  char* end = nullptr;
  v = std::strtod(c, &end);
  return end != c && trim(end).empty();
}

inline bool full_parse_long(const std::string& t, long& v) {
  const char* c = t.c_str();
  char* end = nullptr;
  v = std::strtol(c, &end, 10);
  return end != c && trim(end).empty();
}

}  // namespace prm_detail

inline void DataStorage::read_data(const std::string& filename) {
  using namespace prm_detail;
  std::ifstream f(filename);
  if (!f) throw std::runtime_error("Could not open file " + filename);  // ExcFileNotOpen (:138-139)
  struct Entry {
    std::function<bool(const std::string&)> set;
  };
  auto dbl = [](double& dst, bool (*ok)(double)) {
    return [&dst, ok](const std::string& t) {
      double v;
      if (!full_parse_double(t, v) || !ok(v)) return false;
      dst = v;
      return true;
    };
  };
  auto lng = [](auto& dst, long lo, long hi) {
    return [&dst, lo, hi](const std::string& t) {
      long v;
      if (!full_parse_long(t, v) || v < lo || v > hi) return false;
      dst = static_cast<std::remove_reference_t<decltype(dst)>>(v);
      return true;
    };
  };
  auto nonneg = [](double v) { return v >= 0.0; };
  auto pos = [](double v) { return v > 0.0; };
  auto any = [](double) { return true; };
  const long big = 2147483647L;
  std::map<std::pair<std::string, std::string>, Entry> E = {
      {{"Physical data", "initial_time"}, {dbl(initial_time, nonneg)}},
      {{"Physical data", "final_time"}, {dbl(final_time, nonneg)}},
      {{"Physical data", "Mach"}, {dbl(Mach, nonneg)}},
      {{"Physical data", "Froude"}, {dbl(Froude, nonneg)}},
      {{"Time step data", "dt"}, {dbl(dt, nonneg)}},
      {{"Space discretization", "n_of_refines"}, {lng(n_global_refines, 0, 15)}},
      {{"Data solve", "max_iterations"}, {lng(max_iterations, 1, 30000)}},
      {{"Data solve", "eps"}, {dbl(eps, nonneg)}},
      {{"", "verbose"},
       {[this](const std::string& text) {
         std::string t = text;
         for (char& ch : t) ch = (char)std::tolower((unsigned char)ch);
         if (t == "true" || t == "1" || t == "yes") verbose = true;
         else if (t == "false" || t == "0" || t == "no") verbose = false;
         else return false;
         return true;
       }}},
      {{"", "output_interval"}, {lng(output_interval, 1, big)}},
      {{"", "saving directory"},
       {[this](const std::string& t) {
         dir = t;
         return true;
       }}},
      {{"Device", "degree"}, {lng(degree, 1, 3)}},
      {{"Device", "nx"}, {lng(nx, 1, big)}},
      {{"Device", "nz"}, {lng(nz, 1, big)}},
      {{"Device", "Lx"}, {dbl(Lx, pos)}},
      {{"Device", "H"}, {dbl(H, pos)}},
      {{"Device", "hill_height"}, {dbl(hill_height, nonneg)}},
      {{"Device", "hill_half_width"}, {dbl(hill_half_width, pos)}},
      {{"Device", "hill_center"}, {dbl(hill_center, any)}},
      {{"Device", "precision"},
       {[this](const std::string& t) {
         if (t != "fp64" && t != "mixed" && t != "fp64_direct") return false;
         precision = t;
         return true;
       }}},
      {{"Device", "gravity"}, {dbl(gravity, pos)}},
      {{"Device", "wind"}, {dbl(wind, any)}},
      {{"Device", "brunt_vaisala"}, {dbl(brunt_vaisala, pos)}},
      {{"Device", "seed"}, {lng(seed, 0, 9223372036854775807L)}},
  };
  std::vector<std::string> stack;
  std::string raw;
  int ln = 0;
  auto where = [&]() { return filename + ":" + std::to_string(ln) + ": "; };
  while (std::getline(f, raw)) {
    ++ln;
    const std::string line = trim(raw.substr(0, raw.find('#')));
    if (line.empty()) continue;
    if (line.rfind("subsection ", 0) == 0) {
      stack.push_back(trim(line.substr(11)));
    } else if (line == "end") {
      if (stack.empty()) throw std::runtime_error(where() + "'end' without subsection");
      stack.pop_back();
    } else if (line.rfind("set ", 0) == 0) {
      const size_t eq = line.find('=');
      if (eq == std::string::npos) throw std::runtime_error(where() + "expected 'set key = value'");
      const std::string key = trim(line.substr(4, eq - 4)), val = trim(line.substr(eq + 1));
      const std::string sec = stack.empty() ? "" : stack.back();
      auto it = E.find({sec, key});
      if (it == E.end()) throw std::runtime_error(where() + "unknown entry '" + key + "' in subsection '" + sec + "'");
      if (!it->second.set(val))
        throw std::runtime_error(where() + "value '" + val + "' of '" + key + "' does not match its pattern");
    } else {
      throw std::runtime_error(where() + "cannot parse '" + line + "'");
    }
  }
  if (!stack.empty()) throw std::runtime_error(filename + ": unterminated subsection '" + stack.back() + "'");
  // EulerSolver constructor check (Atmospheric_Flow.cc:256, T = final_time at :220)
  if (dt <= 0.0 || dt > 0.5 * final_time) {
    std::ostringstream o;
    o << "The time step " << dt << " is out of range (0, " << 0.5 * final_time << "]";
    throw std::runtime_error(o.str());
  }
}

inline std::string DataStorage::dump() const {
  std::ostringstream o;
  o.precision(17);
  o << "initial_time = " << initial_time << "\nfinal_time = " << final_time << "\nMach = " << Mach
    << "\nFroude = " << Froude << "\ndt = " << dt << "\nn_global_refines = " << n_global_refines
    << "\nmax_iterations = " << max_iterations << "\neps = " << eps << "\nverbose = " << (verbose ? 1 : 0)
    << "\noutput_interval = " << output_interval << "\ndir = " << dir << "\ndegree = " << degree << "\nnx = " << nx
    << "\nnz = " << nz << "\nLx = " << Lx << "\nH = " << H << "\nhill_height = " << hill_height
    << "\nhill_half_width = " << hill_half_width << "\nhill_center = " << hill_center
    << "\nprecision = " << precision << "\ngravity = " << gravity << "\nwind = " << wind
    << "\nbrunt_vaisala = " << brunt_vaisala << "\nseed = " << seed << "\n";
  return o.str();
}

}  // namespace atmo
