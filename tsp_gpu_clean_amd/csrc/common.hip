// Error state and probes shared by every tspmi entry point.
#include <stdarg.h>
#include <stdio.h>

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
