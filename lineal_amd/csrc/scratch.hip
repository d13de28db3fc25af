// Device scratch of the entry points that allocate nothing
// (lgpu_op_ssor_noalloc).  The array is a file-scope device variable: every
// GPU the library runs on holds its own copy, at its own address, until the
// process ends.  In host code the name `scratch` is only a host-side stand-in
// for the symbol, not a device address, so host code never uses it directly:
// scratch_init() looks up each device's address once, with that device
// current, and every launch takes that address.
#include <mutex>

#include "internal.hpp"

static __device__ __align__(256) double scratch[lgpu::kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_ptr[kMaxDevices] = {};

}  // namespace

namespace lgpu {

void scratch_init(int device) {
  LGPU_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_ptr[device]) return;
  int cur = -1;
  LGPU_HIP_CHECK(hipGetDevice(&cur));
  LGPU_REQUIRE(cur == device, "scratch_init needs its device current");
  void* p = nullptr;
  LGPU_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_ptr[device] = p;
}

double* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_mutex);
  void* p = device >= 0 && device < kMaxDevices ? g_ptr[device] : nullptr;
  LGPU_REQUIRE(p != nullptr, "the device scratch of this device was not set up");
  return static_cast<double*>(p);
}

}  // namespace lgpu
