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

// Calibration probe: every thread of a 256-thread workgroup issues `iters` x 8 independent 4-byte
// loads at pseudo-random offsets inside one 4 KB row of `rows` (row blockIdx.x % nrows), the access
// shape of the 2-opt diagonal scan's distance gathers.  Offsets are masked into the row, so every
// access stays inside the rows_count x 1024 buffer.
__global__ __launch_bounds__(256) void probe_gather_kernel(const int32_t* __restrict__ rows, int nrows, int iters,
                                                           int32_t* __restrict__ out) {
  const int32_t* __restrict__ row = rows + (size_t)(blockIdx.x % (unsigned)nrows) * 1024;
  uint32_t h = (blockIdx.x * 256u + threadIdx.x) * 2654435761u + 12345u;
  int acc = 0;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      h = h * 1664525u + 1013904223u;  // Numerical Recipes LCG
      acc += row[h >> 22];
    }
  }
  if (acc == 0x7fffffff) out[0] = acc;  // keeps the loads live; practically never taken
}

}  // namespace tspmi

extern "C" {

int tspmi_probe_gather(const int32_t* rows, int nrows, int blocks, int iters, int32_t* out, void* stream) {
  TSPMI_CHECK_ARG(rows != nullptr && out != nullptr, "tspmi_probe_gather: null pointer");
  TSPMI_CHECK_ARG(nrows >= 1 && blocks >= 1 && iters >= 0, "tspmi_probe_gather: bad shape");
  hipLaunchKernelGGL(tspmi::probe_gather_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, rows, nrows, iters,
                     out);
  TSPMI_CHECK_LAUNCH("tspmi_probe_gather launch");
  return TSPMI_OK;
}

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
