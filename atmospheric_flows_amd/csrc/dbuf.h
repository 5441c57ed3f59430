// Device buffers owned by an engine, with per-thread accounting of the
// bytes they request (engine constructors measure their own footprint as
// the difference around their allocations).
#pragma once
#include <stdexcept>
#include <string>

#include "common.h"

namespace atmo {

struct NoConvergence : std::runtime_error {
  explicit NoConvergence(const std::string& s) : std::runtime_error(s) {}
};

// Device bytes requested by the DBufs of the calling thread.
extern thread_local size_t tl_dev_bytes;

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void alloc(size_t n_) {
    release();
    if (n_) ATMO_HIP_CHECK(hipMalloc(&p, n_ * sizeof(T)));
    n = n_;
    tl_dev_bytes += n * sizeof(T);
  }
  void zero(hipStream_t s) {
    if (n) ATMO_HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  void release() {
    if (p) (void)hipFree(p);
    tl_dev_bytes -= n * sizeof(T);
    p = nullptr;
    n = 0;
  }
};

}  // namespace atmo
