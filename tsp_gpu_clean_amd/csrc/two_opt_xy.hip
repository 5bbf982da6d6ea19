// 2-opt on EUC_2D instances with distances computed from the coordinates (gfx950).
//
// Same scan order, delta, tie-break, acceptance and move as the matrix kernels of two_opt.hip
// (reference src/algorithms/strategies/local_search/two_opt.py:158-201, 237-245, 347-367,
// 386-440), and the same distance values: every distance is euc2d_nint(), the function that
// builds the matrix.  What changes is where a distance comes from.  A matrix gather is a 4-byte
// random read from a row of 4n bytes, which stays on chip at n = 1000 but goes to HBM once the
// matrix outgrows the caches (400 MB at n = 10000); here it is a 16-byte read from the 16n-byte
// coordinate array (160 KB at n = 10000, L2-resident) plus about twenty FP64 instructions.  No
// n x n matrix is read at all, so the kernels also run where none is built.
#include "two_opt_dev.hpp"

namespace tspmi {

// Diagonal lane mapping with register carry, as scan_full_best_diag in two_opt.hip: lane (wave w,
// lane L) owns j = i + 2 + 64 w + L + q * BLOCK of every row i; the slot's second distance of row
// i, d(t_{i+1}, t_{j+1}), is its first distance of row i+1, and the next row's distances are
// computed before the current row's deltas (one-row pipeline).  Out-of-range slots read clamped
// positions and are masked by the bound test.
template <int BLOCK, int QMAX>
__device__ __forceinline__ uint64_t scan_full_best_xy(const int32_t* s_tour, const int32_t* s_edge, int n,
                                                      const double2* __restrict__ xy, int wid, int lane) {
  const int rows = n - 2;
  const int jq0 = 2 + wid * kWave + lane;
  int carry[QMAX], gcur[QMAX], bd[QMAX], bi[QMAX];
  {
    const double2 r0 = xy[__builtin_amdgcn_readfirstlane(s_tour[0])];
    const double2 r1 = xy[__builtin_amdgcn_readfirstlane(s_tour[1])];
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int j = jq0 + q * BLOCK;
      const double2 c = xy[s_tour[min(j, n)]];
      const double2 d = xy[s_tour[min(j + 1, n)]];
      carry[q] = euc2d_nint(r0.x, r0.y, c.x, c.y);
      gcur[q] = euc2d_nint(r1.x, r1.y, d.x, d.y);
      bd[q] = 0;
      bi[q] = 0;
    }
  }
  for (int i = 0; i < rows; ++i) {
    if (i + 2 + kWave * wid >= n) break;  // every column of this wave lies past the tour
    const double2 rn = xy[__builtin_amdgcn_readfirstlane(s_tour[i + 2])];
    const int eab = s_edge[i];
    const int lim = (i == 0) ? n - 1 : n;
    // Per slot: the next row's distance, then this row's delta; no per-slot temporaries outlive
    // the slot, so the live state stays at four registers per slot.
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int j = i + jq0 + q * BLOCK;
      const double2 p = xy[s_tour[min(j + 2, n)]];
      const int gnext = euc2d_nint(rn.x, rn.y, p.x, p.y);
      const int delta = carry[q] + gcur[q] - eab - s_edge[min(j, n - 1)];
      const bool take = (j < lim) & (delta < bd[q]);
      bd[q] = take ? delta : bd[q];
      bi[q] = take ? i : bi[q];
      carry[q] = gcur[q];
      gcur[q] = gnext;
      // Keep the scheduler from hoisting every slot's loads and FP64 temporaries at once (that
      // spills at 1024 threads); four slots in flight per group.
      if ((q & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
  uint64_t best = kNoMove;
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    if (bd[q] < 0) {
      const uint32_t r = (uint32_t)bi[q];
      const uint64_t key = pack_key(bd[q], r * (uint32_t)n + r + (uint32_t)(jq0 + q * BLOCK));
      best = key < best ? key : best;
    }
  }
  return best;
}

template <int BLOCK, int QMAX>
__global__ __launch_bounds__(BLOCK) void two_opt_xy_kernel(int32_t* __restrict__ tours, int n,
                                                           const double2* __restrict__ xy, int max_iter,
                                                           int64_t* __restrict__ stats) {
  extern __shared__ int32_t lds[];
  __shared__ uint64_t red[BLOCK / kWave];
  int32_t* s_tour = lds;          // n + 1 (s_tour[n] mirrors s_tour[0])
  int32_t* s_edge = lds + n + 1;  // n
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = wave_id();
  const CoordDist dist{xy};
  int32_t* tour = tours + (size_t)blockIdx.x * n;
  for (int x = tid; x < n; x += BLOCK) {
    const int v = clamp_id(tour[x], n);
    s_tour[x] = v;
    if (x == 0) s_tour[n] = v;
  }
  __syncthreads();
  for (int x = tid; x < n; x += BLOCK) s_edge[x] = dist(s_tour[x], s_tour[x + 1]);
  __syncthreads();
  int moves = 0, passes = 0;
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    uint64_t best = scan_full_best_xy<BLOCK, QMAX>(s_tour, s_edge, n, xy, wid, lane);
    best = block_min_u64(best, red);
    if (best == kNoMove) break;
    const uint32_t idx = key_index(best);
    const int i = (int)(idx / (uint32_t)n);
    const int j = (int)(idx - (uint32_t)i * (uint32_t)n);
    __syncthreads();
    apply_move(s_tour, s_edge, nullptr, i, j, n, dist);
    ++moves;
  }
  for (int x = tid; x < n; x += BLOCK) tour[x] = s_tour[x];
  if (stats != nullptr && tid == 0) {
    atomicAdd((unsigned long long*)&stats[0], (unsigned long long)((int64_t)passes * n * (n - 3) / 2));
    atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
    atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
  }
}

// XYD: CoordDist (float64 coordinates) or IntCoordDist (int32 coordinates, the integer form).
// XY_LDS (IntCoordDist only): rec_two_opt_tour, with the coordinates and node records in LDS (its
// comment); otherwise cand_two_opt_tour reading the coordinates from global memory.
template <int BLOCK, int KP, int UNROLL, int KL, class XYD, bool XY_LDS = false>
__global__ __launch_bounds__(BLOCK) void two_opt_cand_xy_kernel(int32_t* __restrict__ tours, int n, XYD dist,
                                                                const int2* __restrict__ ct, int k, int max_iter,
                                                                int64_t* __restrict__ stats,
                                                                double* __restrict__ cost,
                                                                int64_t* __restrict__ pass_log, int log_moves) {
  extern __shared__ __align__(16) int32_t lds[];
  __shared__ uint64_t red[BLOCK / kWave];
  __shared__ int64_t red64[BLOCK / kWave];
  double* const c = cost != nullptr ? cost + blockIdx.x : nullptr;
  int64_t* const pl = pass_log != nullptr ? pass_log + (size_t)blockIdx.x * 2 * (log_moves + 1) : nullptr;
  if constexpr (XY_LDS)
    rec_two_opt_tour<BLOCK, KP, UNROLL, KL>(tours + (size_t)blockIdx.x * n, n, dist.xy, ct, k, max_iter, stats, lds,
                                            red, red64, c, pl, log_moves);
  else
    cand_two_opt_tour<BLOCK, KP, UNROLL, KL>(tours + (size_t)blockIdx.x * n, n, dist, ct, k, max_iter, stats, lds,
                                             red, red64, c, pl, log_moves);
}

}  // namespace tspmi

using namespace tspmi;

extern "C" int tspmi_two_opt_euc2d(int32_t* tours, int P, int n, const double* xy, int max_iter, int64_t* stats,
                                   void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_two_opt_euc2d: negative shape");
  if (P == 0 || n < 4 || max_iter == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(tours && xy, "tspmi_two_opt_euc2d: null pointer");
  TSPMI_CHECK_ARG(((uintptr_t)xy & 15) == 0, "tspmi_two_opt_euc2d: xy must be 16-byte aligned");
  TSPMI_CHECK_ARG((long)n * (long)n < (1L << 32), "tspmi_two_opt_euc2d: n=%d too large for 32-bit scan indices", n);
  const size_t lds = (size_t)(2 * n + 1) * sizeof(int32_t);
  TSPMI_CHECK_ARG(lds + 1024 <= (size_t)kMaxLds, "tspmi_two_opt_euc2d: n=%d exceeds the LDS limit", n);
  const int cols = n - 2;  // largest j - i + 1 a slot must reach
  // Eight slots per lane at 1024 threads is the register limit (more spill), hence n <= 8194.
  TSPMI_CHECK_ARG(cols <= 8 * 1024 + 1, "tspmi_two_opt_euc2d: n=%d above 8194; use the matrix kernel", n);
  hipStream_t s = (hipStream_t)stream;
  const double2* p = reinterpret_cast<const double2*>(xy);
  const bool few = P <= device_cu_count();
#define TSPMI_XY(BL, Q)                                                                                        \
  hipLaunchKernelGGL((two_opt_xy_kernel<BL, Q>), dim3(P), dim3(BL), lds, s, tours, n, p, max_iter, stats)
  if (!few && cols <= 4 * 256 + 1) TSPMI_XY(256, 4);
  else if (!few && cols <= 8 * 256 + 1) TSPMI_XY(256, 8);
  else if (cols <= 2 * 1024 + 1) TSPMI_XY(1024, 2);
  else if (cols <= 4 * 1024 + 1) TSPMI_XY(1024, 4);
  else TSPMI_XY(1024, 8);
#undef TSPMI_XY
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_euc2d launch");
  return TSPMI_OK;
}

namespace {

// `xy` is float64 (n, 2) (16-byte aligned), or, with int_xy, int32 (n, 2) (8-byte aligned).
int packed_euc2d(int32_t* tours, int P, int n, const void* xy, bool int_xy, const int32_t* table, int k,
                 int max_iter, int64_t* stats, double* cost, int64_t* pass_log, int log_moves, void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_two_opt_packed_euc2d: negative shape");
  TSPMI_CHECK_ARG(k >= 1 && k <= 32, "tspmi_two_opt_packed_euc2d: k=%d outside [1, 32]", k);
  if (P == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(cost == nullptr || (n >= 4 && max_iter != 0),
                  "tspmi_two_opt_packed_euc2d: costs need n >= 4 and max_iter != 0");
  if (n < 4 || max_iter == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(tours && xy && table, "tspmi_two_opt_packed_euc2d: null pointer");
  TSPMI_CHECK_ARG(((uintptr_t)xy & (int_xy ? 7 : 15)) == 0 && ((uintptr_t)table & 7) == 0,
                  "tspmi_two_opt_packed_euc2d: xy must be %d-byte and table 8-byte aligned", int_xy ? 8 : 16);
  TSPMI_CHECK_ARG((long)(n - 2) * k < (1L << 31), "tspmi_two_opt_packed_euc2d: (n-2)*k too large");
  const size_t lds = (size_t)(3 * n + 1) * sizeof(int32_t);
  TSPMI_CHECK_ARG(lds + 1024 <= (size_t)kMaxLds, "tspmi_two_opt_packed_euc2d: n=%d exceeds the LDS limit", n);
  const size_t lds_xy = (size_t)rec_lds_words(n) * sizeof(int32_t);
  const bool xy_lds = int_xy && lds_xy + 1024 <= (size_t)kMaxLds;
  hipStream_t s = (hipStream_t)stream;
  const int2* ct = reinterpret_cast<const int2*>(table);
  // One tour per CU: 16 waves each; up to four per CU: 8 waves (measured at P = 1024, n = 1000 and
  // 2000, k = 20: 0.90 / 1.73 ms per 20-pass launch at 512 threads vs 0.97 / 1.93 at 256); more: 4.
  const int cus = device_cu_count();
  const int block = P <= cus ? 1024 : (P <= 4 * cus ? 512 : 256);
#define TSPMI_CXY(BL, KPV, KLV)                                                                                      \
  if (xy_lds)                                                                                                        \
    hipLaunchKernelGGL((two_opt_cand_xy_kernel<BL, KPV, 4, KLV, IntCoordDist, true>), dim3(P), dim3(BL), lds_xy, s,  \
                       tours, n, IntCoordDist{reinterpret_cast<const int2*>(xy)}, ct, k, max_iter, stats, cost,      \
                       pass_log, log_moves);                                                                         \
  else if (int_xy)                                                                                                   \
    hipLaunchKernelGGL((two_opt_cand_xy_kernel<BL, KPV, 4, KLV, IntCoordDist>), dim3(P), dim3(BL), lds, s, tours, n, \
                       IntCoordDist{reinterpret_cast<const int2*>(xy)}, ct, k, max_iter, stats, cost, pass_log,      \
                       log_moves);                                                                                   \
  else                                                                                                               \
    hipLaunchKernelGGL((two_opt_cand_xy_kernel<BL, KPV, 4, KLV, CoordDist>), dim3(P), dim3(BL), lds, s, tours, n,    \
                       CoordDist{reinterpret_cast<const double2*>(xy)}, ct, k, max_iter, stats, cost, pass_log,      \
                       log_moves)
  if (block == 1024) {
    TSPMI_CAND_BY_K(TSPMI_CXY, 1024, k)
  } else if (block == 512) {
    TSPMI_CAND_BY_K(TSPMI_CXY, 512, k)
  } else {
    TSPMI_CAND_BY_K(TSPMI_CXY, 256, k)
  }
#undef TSPMI_CXY
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_packed_euc2d launch");
  return TSPMI_OK;
}

}  // namespace

extern "C" int tspmi_two_opt_packed_euc2d_cost(int32_t* tours, int P, int n, const double* xy, const int32_t* table,
                                               int k, int max_iter, int64_t* stats, double* cost, void* stream) {
  return packed_euc2d(tours, P, n, xy, false, table, k, max_iter, stats, cost, nullptr, 0, stream);
}

extern "C" int tspmi_two_opt_packed_euc2d_log(int32_t* tours, int P, int n, const double* xy, const int32_t* table,
                                              int k, int max_iter, int64_t* stats, double* cost, int64_t* pass_log,
                                              int log_moves, void* stream) {
  TSPMI_CHECK_ARG(pass_log != nullptr && log_moves >= 0, "tspmi_two_opt_packed_euc2d_log: bad pass log");
  return packed_euc2d(tours, P, n, xy, false, table, k, max_iter, stats, cost, pass_log, log_moves, stream);
}

extern "C" int tspmi_two_opt_packed_euc2d_i32(int32_t* tours, int P, int n, const int32_t* ixy, const int32_t* table,
                                              int k, int max_iter, int64_t* stats, double* cost, int64_t* pass_log,
                                              int log_moves, void* stream) {
  TSPMI_CHECK_ARG(pass_log == nullptr || log_moves >= 0, "tspmi_two_opt_packed_euc2d_i32: bad pass log");
  return packed_euc2d(tours, P, n, ixy, true, table, k, max_iter, stats, cost, pass_log, pass_log ? log_moves : 0,
                      stream);
}

extern "C" int tspmi_wall_clock_khz(int* khz) {
  TSPMI_CHECK_ARG(khz != nullptr, "tspmi_wall_clock_khz: null pointer");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, dev);
  if (e != hipSuccess) return hip_fail(e, "tspmi_wall_clock_khz");
  return TSPMI_OK;
}

extern "C" int tspmi_two_opt_packed_euc2d(int32_t* tours, int P, int n, const double* xy, const int32_t* table, int k,
                                          int max_iter, int64_t* stats, void* stream) {
  return tspmi_two_opt_packed_euc2d_cost(tours, P, n, xy, table, k, max_iter, stats, nullptr, stream);
}
