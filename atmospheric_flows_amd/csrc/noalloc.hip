// Scalar product on caller-owned device memory that allocates nothing:
// atmo_dot_noalloc (include/atmo_hip.h). It runs the fixed-order reduction of
// atmo_vec_dot (one partial per block, then one finalize block summing them in
// block order into the caller's result) with the block partials in `scratch`,
// a file-scope device array of a fixed size, instead of a context's reduction
// buffer. Each GPU holds its own copy of the array until the process ends; its
// address on a device is looked up once, with that device current, and only
// that address is handed to kernels (the array's name in host code is a
// host-side stand-in, not a device address). Every call on one GPU shares that
// GPU's array: calls that overlap in time on one GPU race on it.
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/atmo_hip.h"
#include "krylov.h"

// error bookkeeping of the C ABI (capi.cpp): the message is read with atmo_last_error(NULL)
atmo_status atmo_guarded_noctx(const std::function<atmo_status()>& f);

namespace atmo {
namespace {

// the block partials of the fixed-order sum (the finalize block writes the result to the caller's memory)
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)RED_MAX_BLOCKS;
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

std::mutex g_mu;
double* g_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up on the first call on that device.
double* scratch_ptr() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    throw std::runtime_error("atmo_dot_noalloc: no usable current HIP device");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw DeviceError(std::string("atmo_dot_noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_ptr[dev] = (double*)p;
  }
  return g_ptr[dev];
}

}  // namespace
}  // namespace atmo

extern "C" atmo_status atmo_dot_noalloc(const double* x, const double* y, size_t n, void* stream, double* result) {
  return atmo_guarded_noctx([&]() -> atmo_status {
    using namespace atmo;
    if (!x || !y || !result) throw std::invalid_argument("atmo_dot_noalloc: null pointer");
    if (n > (size_t)ATMO_DOT_NOALLOC_MAX_N) throw std::invalid_argument("atmo_dot_noalloc: n > ATMO_DOT_NOALLOC_MAX_N");
    RedBuf rb;
    rb.gen = scratch_ptr();  // vec_dot uses only the general partials slot
    vec_dot((long)n, x, y, result, false, rb, (hipStream_t)stream);
    return ATMO_OK;
  });
}
