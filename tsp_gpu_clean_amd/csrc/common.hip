// Error state and probes shared by every tspmi entry point.
#include <stdarg.h>
#include <stdio.h>

#include <atomic>

#include "common.hpp"

namespace tspmi {

static thread_local char g_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t err, const char* what) {
  set_error("%s: %s", what, hipGetErrorString(err));
  return TSPMI_EHIP;
}

int device_cu_count() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  int v = cache[dev].load(std::memory_order_relaxed);
  if (v <= 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cache[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

}  // namespace tspmi

extern "C" {

int tspmi_version(void) { return 1; }

const char* tspmi_last_error(void) { return tspmi::g_error; }

int tspmi_device_count(int* count) {
  TSPMI_CHECK_ARG(count != nullptr, "tspmi_device_count: count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess || c <= 0) {
    *count = 0;
    tspmi::set_error("no HIP device visible");
    return TSPMI_ENODEV;
  }
  *count = c;
  return TSPMI_OK;
}

}  // extern "C"
