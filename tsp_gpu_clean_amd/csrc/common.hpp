// Shared helpers for the tspmi HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tspmi.h"

namespace tspmi {

constexpr int kWave = 64;

// Thread-local last-error text, set by every failing entry point.
void set_error(const char* fmt, ...);

// Map a HIP status to the library status, recording the message.
int hip_fail(hipError_t err, const char* what);

// Compute units of the current device (cached per device; 256 if the query fails).
int device_cu_count();

#define TSPMI_CHECK_ARG(cond, ...)      \
  do {                                  \
    if (!(cond)) {                      \
      ::tspmi::set_error(__VA_ARGS__);  \
      return TSPMI_EINVAL;              \
    }                                   \
  } while (0)

#define TSPMI_CHECK_LAUNCH(what)                          \
  do {                                                    \
    hipError_t _e = hipGetLastError();                    \
    if (_e != hipSuccess) return ::tspmi::hip_fail(_e, what); \
  } while (0)

// Largest dynamic LDS a single workgroup may request on CDNA4 (160 KiB per CU).
constexpr int kMaxLds = 160 * 1024;

// ---- wave64 reductions -------------------------------------------------------

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    uint64_t o = __shfl_xor(v, m, kWave);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

// Workgroup-wide minimum of a 64-bit key; every thread gets the result.
// `scratch` must hold blockDim.x / 64 entries.
__device__ __forceinline__ uint64_t block_min_u64(uint64_t v, uint64_t* scratch) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  v = wave_min_u64(v);
  __syncthreads();  // scratch may still be read from a previous call
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  uint64_t r = scratch[0];
  for (int w = 1; w < nw; ++w) r = scratch[w] < r ? scratch[w] : r;
  return r;
}

__device__ __forceinline__ int64_t block_sum_i64(int64_t v, int64_t* scratch) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  v = wave_sum_i64(v);
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  int64_t r = 0;
  for (int w = 0; w < nw; ++w) r += scratch[w];
  return r;
}

// Order-preserving packing of (signed delta, scan index): the unsigned minimum picks the
// most negative delta and, among equal deltas, the lowest scan index.
__device__ __forceinline__ uint64_t pack_key(int delta, uint32_t idx) {
  return ((uint64_t)((uint32_t)delta ^ 0x80000000u) << 32) | (uint64_t)idx;
}
__device__ __forceinline__ int key_delta(uint64_t key) { return (int)((uint32_t)(key >> 32) ^ 0x80000000u); }
__device__ __forceinline__ uint32_t key_index(uint64_t key) { return (uint32_t)key; }

constexpr uint64_t kNoMove = ~0ull;

// Node ids outside [0, limit) are an invalid input; clamping keeps every access in bounds.
__device__ __forceinline__ int clamp_id(int v, int limit) { return (unsigned)v < (unsigned)limit ? v : 0; }

// EUC_2D TSPLIB nint with FP64 kept for the sqrt-and-round step only:
// d = (int)(sqrt(dx*dx + dy*dy) + 0.5), products and sum rounded separately (no FMA
// contraction), as NumPy evaluates src/distances/matrix.py:84-86.  Every kernel that needs an
// EUC_2D distance calls this one function, so a distance computed on the fly is bit-identical to
// the matrix entry built by tspmi_dist_euc2d.
__device__ __forceinline__ int euc2d_nint(double xu, double yu, double xv, double yv) {
#pragma clang fp contract(off)
  const double dx = xu - xv;
  const double dy = yu - yv;
  const double s = dx * dx + dy * dy;
  return (int)(__builtin_sqrt(s) + 0.5);
}

}  // namespace tspmi
