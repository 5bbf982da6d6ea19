// Shared helpers for the tspmi HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tspmi.h"

namespace tspmi {

constexpr int kWave = 64;
// Accelerator complex dies (XCDs) of an MI355X; each has its own L2 and receives workgroups in turn.
constexpr int kXcds = 8;

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

// Index of the calling wave in its workgroup, in a scalar register: every lane of a wave holds
// the same value, and saying so lets loops bounded by it run on scalar counters.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave); }

// ---- wave64 reductions -------------------------------------------------------

// One DPP step of a 64-bit minimum: lanes of the rows in `row_mask` take min(v, v of the lane that
// CTRL selects); other lanes keep v.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp_min_u64(uint64_t v) {
  const int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
  const uint32_t olo = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
  const uint32_t ohi = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
  const uint64_t o = ((uint64_t)ohi << 32) | olo;
  return o < v ? o : v;
}

// Minimum over the 64 lanes of a wave, returned to every lane; the whole wave must be active.
// Data-parallel-primitive moves instead of LDS permutes: quad swaps (xor 1, xor 2), half-row
// and row mirrors (each row of 16 then holds its minimum), then row broadcasts 15 and 31 carry
// rows 0-2 into row 3, and lane 63 is read out.
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
  v = dpp_min_u64<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
  v = dpp_min_u64<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
  v = dpp_min_u64<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_min_u64<0x140, 0xf>(v);  // row_mirror
  v = dpp_min_u64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_min_u64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

// The same minimum for a 32-bit signed value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_min_i32(int v) {
  return min(v, __builtin_amdgcn_update_dpp(v, v, CTRL, ROW_MASK, 0xf, false));
}
__device__ __forceinline__ int wave_min_i32(int v) {
  v = dpp_min_i32<0xb1, 0xf>(v);
  v = dpp_min_i32<0x4e, 0xf>(v);
  v = dpp_min_i32<0x141, 0xf>(v);
  v = dpp_min_i32<0x140, 0xf>(v);
  v = dpp_min_i32<0x142, 0xa>(v);
  v = dpp_min_i32<0x143, 0xc>(v);
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
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
__device__ __forceinline__ uint32_t key_index(uint64_t key) { return (uint32_t)key; }

constexpr uint64_t kNoMove = ~0ull;

// Unsigned bits of a double that order as NumPy's sorts order doubles: -0.0 and +0.0 equal, and
// every NaN (either sign, any payload) after +inf and equal to every other NaN, so that equal
// keys broken by index give np.argsort(kind="stable") (NaNs last, in index order).
__device__ __forceinline__ uint64_t orderable_bits(double v) {
  if (v != v) return ~0ull;
  if (v == 0.0) v = 0.0;
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Key of np.argmin over doubles: a NaN propagates, so the first NaN is the minimum; other values
// order as orderable_bits (which never returns 0 for a number: -inf maps to 2^52 - 1).
__device__ __forceinline__ uint64_t argmin_bits(double v) { return v != v ? 0ull : orderable_bits(v); }

// v unchanged, in a VGPR, through an empty asm statement: the optimiser cannot see through it, so
// it neither reassociates the arithmetic that produced v with what consumes it nor relates two
// values laundered separately.  Emits no instruction.
__device__ __forceinline__ int opaque_i32(int v) {
  asm("" : "+v"(v));
  return v;
}

// Index of (x, y) in [0, 2^16)^2 along the Hilbert curve of order 16.
__device__ __forceinline__ uint32_t hilbert_index16(uint32_t x, uint32_t y) {
  constexpr uint32_t kSide = 1u << 16;
  uint32_t d = 0;
  for (uint32_t s = kSide >> 1; s > 0; s >>= 1) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = kSide - 1 - x;
        y = kSide - 1 - y;
      }
      const uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

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

// Largest per-axis coordinate span for which euc2d_nint_i32 is exact.
constexpr int kIntCoordSpan = 1 << 20;

// euc2d_nint for integer coordinates whose per-axis differences are at most 2^20 in magnitude,
// with the same result.  Why the two agree: s = dx^2 + dy^2 <= 2^41 is an exact integer in FP64,
// and a square root of an integer lies at least 0.25 / (2 sqrt(s) + 1) > 8e-8 from every
// half-integer, far more than the < 1e-9 that the FP64 square root and the + 0.5 can round it
// by, so the FP64 expression returns the exact nint(sqrt(s)): the M with M^2 - M < s <= M^2 + M
// (M = 0 for s = 0).  Here: r = sqrt(dx^2 + dy^2) in FP32 (two roundings of the sum, one of the
// hardware square root: relative error below 2^-21) is within 0.3 of sqrt(s) <= 2^20.5, so
// m = trunc(r) is M - 1 or M.  For m = M - 1, m^2 + m = M^2 - M < s, so t = m^2 + m - s < 0; for
// m = M, t >= 0; so m - (t >> 31) is M.  t (below 2^22 in magnitude) is exact when formed modulo
// 2^32 from 24-bit products.
// Fourteen 32-bit VALU instructions against about twenty-five, some of them FP64, for the FP64 form.
__device__ __forceinline__ int euc2d_nint_i32(int xu, int yu, int xv, int yv) {
  const int dx = xu - xv;
  const int dy = yu - yv;
  const float fx = (float)dx, fy = (float)dy;
  const int m = (int)__builtin_amdgcn_sqrtf(__builtin_fmaf(fx, fx, fy * fy));
  const uint32_t s = (uint32_t)__mul24(dx, dx) + (uint32_t)__mul24(dy, dy);
  const int t = (int)((uint32_t)__umul24((uint32_t)m, (uint32_t)m) + (uint32_t)m - s);
  return m - (t >> 31);
}

// A lower bound of euc2d_nint_i32 that is at most 1 below it, in six VALU instructions, for
// integer coordinates held as FP32 values (exact: each is an integer of magnitude at most 2^24)
// whose per-axis differences are at most 2^20: the differences are then exact too, and the
// truncated FP32 root is the m of euc2d_nint_i32 before its exact correction, M - 1 or M (same
// argument).  For scans that only need to bracket the best delta and then settle the few pairs
// near it with exact distances.
__device__ __forceinline__ int euc2d_floor_f32(float xu, float yu, float xv, float yv) {
  const float fx = xu - xv, fy = yu - yv;
  return (int)__builtin_amdgcn_sqrtf(__builtin_fmaf(fx, fx, fy * fy));
}

// Distance policies: a row-major int32 matrix, or EUC_2D coordinates with the distance computed
// on the fly by the same FP64 function that builds the matrix (bit-identical values).
// Every entry point that reads a matrix through MatrixDist requires n * ld < 2^30, so an element's
// byte offset fits 32 bits and the gather is addressed from the matrix base with one 32-bit offset.
constexpr long kMaxMatrixElems = 1L << 30;

struct MatrixDist {
  const int32_t* __restrict__ D;
  int ld;
  __device__ __forceinline__ int operator()(int u, int v) const {
    const uint32_t off = ((uint32_t)u * (uint32_t)ld + (uint32_t)v) << 2;
    return *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(D) + off);
  }
};

struct CoordDist {
  const double2* __restrict__ xy;
  __device__ __forceinline__ int operator()(int u, int v) const {
    const double2 a = xy[u], b = xy[v];
    return euc2d_nint(a.x, a.y, b.x, b.y);
  }
};

// The same distances from int32 coordinates by the integer form (every per-axis difference at
// most kIntCoordSpan, which the caller guarantees): half the bytes per coordinate read and about
// half the VALU cycles per distance.
struct IntCoordDist {
  const int2* __restrict__ xy;
  __device__ __forceinline__ int operator()(int u, int v) const {
    const int2 a = xy[u], b = xy[v];
    return euc2d_nint_i32(a.x, a.y, b.x, b.y);
  }
};

// Either coordinate form, chosen per launch (ixy given: the integer form); for the kernels'
// cold paths, where one uniform branch per distance costs nothing that matters.
struct AnyCoordDist {
  const double2* __restrict__ xy;
  const int2* __restrict__ ixy;
  __device__ __forceinline__ int operator()(int u, int v) const {
    return ixy != nullptr ? IntCoordDist{ixy}(u, v) : CoordDist{xy}(u, v);
  }
};


}  // namespace tspmi
