// 2-opt beyond the LDS-resident limits (gfx950): instances whose tour arrays do not fit next to
// each other in a workgroup's 160 KB of LDS (candidate scan above n ~ 13,500, full scan above
// n = 16,386).  The tour, its edge costs and (candidate mode) the inverse positions of each tour
// live in a caller workspace in global memory, one slice per tour, as the reference's candidate
// kernel keeps no tour in shared memory above its tile limit (two_opt.py:75-94, 707-709); the
// slices are read with coalesced loads from L2 and written by the move's reversal.  Scan order,
// delta, tie-break, acceptance and move are those of the LDS kernels (two_opt.hip), so tours and
// counters are identical wherever both run.
//
// Full neighbourhood: the diagonal lane mapping of two_opt.hip with its register carry, run over
// the row offsets j - i in tiles of 1024 * QMAX (one tile per sweep of the rows), the per-slot
// first minimum and the packed (delta, i * n + j) key making the tiles' minimum the scan's.
#include "two_opt_dev.hpp"

namespace tspmi {

constexpr int kLargeBlock = 1024;
constexpr int kLargeWaves = kLargeBlock / kWave;

// Distance rows: row(a) gives the distances from city a to any city v.
struct MatrixRows {
  const int32_t* __restrict__ D;
  int ld;
  struct Row {
    const int32_t* __restrict__ p;
    __device__ __forceinline__ int operator()(int v) const { return row_at(p, v); }
  };
  __device__ __forceinline__ Row row(int a) const { return Row{D + (size_t)a * ld}; }
  __device__ __forceinline__ int operator()(int u, int v) const { return MatrixDist{D, ld}(u, v); }
};

struct CoordRows {
  const double2* __restrict__ xy;
  struct Row {
    double2 p;
    const double2* __restrict__ xy;
    __device__ __forceinline__ int operator()(int v) const {
      const double2 b = xy[v];
      return euc2d_nint(p.x, p.y, b.x, b.y);
    }
  };
  __device__ __forceinline__ Row row(int a) const { return Row{xy[a], xy}; }
  __device__ __forceinline__ int operator()(int u, int v) const { return CoordDist{xy}(u, v); }
};

// Steepest scan of the row offsets j - i in [tb + 2, tb + 2 + 1024 * QMAX): lane L of wave w owns
// j - i = 2 + tb + 64 w + L + 1024 q.  tour has n + 1 entries (tour[n] = tour[0]), edge n.
template <int QMAX, class Rows>
__device__ __forceinline__ uint64_t scan_tile_best(const int32_t* tour, const int32_t* edge, int n, const Rows& rows_of,
                                                   int tb, int wid, int lane) {
  const int rows = n - 2;
  const int jq0 = 2 + tb + wid * kWave + lane;
  int carry[QMAX], gcur[QMAX], bd[QMAX], bi[QMAX];
  {
    const auto r0 = rows_of.row(__builtin_amdgcn_readfirstlane(tour[0]));
    const auto r1 = rows_of.row(__builtin_amdgcn_readfirstlane(tour[1]));
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int j = jq0 + q * kLargeBlock;
      carry[q] = r0(tour[min(j, n)]);
      gcur[q] = r1(tour[min(j + 1, n)]);
      bd[q] = 0;
      bi[q] = 0;
    }
  }
  for (int i = 0; i < rows; ++i) {
    if (i + 2 + tb + kWave * wid >= n) break;  // every offset of this wave lies past row i's end
    const auto rn = rows_of.row(__builtin_amdgcn_readfirstlane(tour[i + 2]));
    const int eab = edge[i];
    const int lim = (i == 0) ? n - 1 : n;
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int j = i + jq0 + q * kLargeBlock;
      const int gnext = rn(tour[min(j + 2, n)]);
      const int delta = carry[q] + gcur[q] - eab - edge[min(j, n - 1)];
      const bool take = (j < lim) & (delta < bd[q]);
      bd[q] = take ? delta : bd[q];
      bi[q] = take ? i : bi[q];
      carry[q] = gcur[q];
      gcur[q] = gnext;
    }
  }
  uint64_t best = kNoMove;
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    if (bd[q] < 0) {
      const uint32_t r = (uint32_t)bi[q];
      const uint64_t key = pack_key(bd[q], r * (uint32_t)n + r + (uint32_t)(jq0 + q * kLargeBlock));
      best = key < best ? key : best;
    }
  }
  return best;
}

template <int QMAX, class Rows>
__global__ __launch_bounds__(kLargeBlock) void two_opt_full_ws_kernel(int32_t* __restrict__ tours, int n, Rows rows_of,
                                                                      int max_iter, int64_t* __restrict__ stats,
                                                                      double* __restrict__ cost, int32_t* ws,
                                                                      long ws_stride) {
  __shared__ uint64_t red[kLargeWaves];
  __shared__ int64_t red64[kLargeWaves];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = wave_id();
  int32_t* tour = tours + (size_t)blockIdx.x * n;
  int32_t* w_tour = ws + (size_t)blockIdx.x * ws_stride;  // n + 1
  int32_t* w_edge = w_tour + n + 1;                        // n
  for (int x = tid; x < n; x += kLargeBlock) {
    const int v = clamp_id(tour[x], n);
    w_tour[x] = v;
    if (x == 0) w_tour[n] = v;
  }
  __syncthreads();
  for (int x = tid; x < n; x += kLargeBlock) w_edge[x] = rows_of(w_tour[x], w_tour[x + 1]);
  __syncthreads();
  const int rows = n - 2;
  int moves = 0, passes = 0;
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    uint64_t best = kNoMove;
    for (int tb = 0; tb < rows; tb += kLargeBlock * QMAX) {
      const uint64_t b = scan_tile_best<QMAX>(w_tour, w_edge, n, rows_of, tb, wid, lane);
      best = b < best ? b : best;
    }
    best = block_min_u64(best, red);
    if (best == kNoMove) break;
    const uint32_t idx = key_index(best);
    const int i = (int)(idx / (uint32_t)n);
    const int j = (int)(idx - (uint32_t)i * (uint32_t)n);
    __syncthreads();
    apply_move(w_tour, w_edge, nullptr, i, j, n, rows_of);
    ++moves;
  }
  for (int x = tid; x < n; x += kLargeBlock) tour[x] = w_tour[x];
  if (cost != nullptr) {
    int64_t part = 0;
    for (int x = tid; x < n; x += kLargeBlock) part += w_edge[x];
    const int64_t total = block_sum_i64(part, red64);
    if (tid == 0) cost[blockIdx.x] = (double)total;
  }
  if (stats != nullptr && tid == 0) {
    atomicAdd((unsigned long long*)&stats[0], (unsigned long long)((int64_t)passes * n * (n - 3) / 2));
    atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
    atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
  }
}

// Candidate scan: cand_two_opt_tour on the tour's workspace slice instead of LDS.
template <int BLOCK, int KP, int KL, class Rows>
__global__ __launch_bounds__(BLOCK) void two_opt_cand_ws_kernel(int32_t* __restrict__ tours, int n, Rows rows_of,
                                                                const int2* __restrict__ ct, int k, int max_iter,
                                                                int64_t* __restrict__ stats,
                                                                double* __restrict__ cost, int32_t* ws,
                                                                long ws_stride) {
  __shared__ uint64_t red[BLOCK / kWave];
  __shared__ int64_t red64[BLOCK / kWave];
  cand_two_opt_tour<BLOCK, KP, 4, KL>(tours + (size_t)blockIdx.x * n, n, rows_of, ct, k, max_iter, stats,
                                      ws + (size_t)blockIdx.x * ws_stride, red, red64,
                                      cost != nullptr ? cost + blockIdx.x : nullptr);
}

}  // namespace tspmi

using namespace tspmi;

namespace {

template <class Rows>
int launch_large(int32_t* tours, int P, int n, const Rows& rows_of, const int2* ct, int k, int max_iter,
                 int64_t* stats, double* cost, int32_t* ws, hipStream_t s) {
  const long stride = 3L * n + 1;
  if (ct == nullptr) {
    const int cols = n - 2;
    if (cols <= 4 * kLargeBlock)
      hipLaunchKernelGGL((two_opt_full_ws_kernel<4, Rows>), dim3(P), dim3(kLargeBlock), 0, s, tours, n, rows_of,
                         max_iter, stats, cost, ws, stride);
    else
      hipLaunchKernelGGL((two_opt_full_ws_kernel<8, Rows>), dim3(P), dim3(kLargeBlock), 0, s, tours, n, rows_of,
                         max_iter, stats, cost, ws, stride);
  } else {
    const int cus = device_cu_count();
    const int block = P <= cus ? 1024 : (P <= 2 * cus ? 512 : 256);
#define TSPMI_LAUNCH_CWS(BL, KPV, KLV)                                                                              \
  hipLaunchKernelGGL((two_opt_cand_ws_kernel<BL, KPV, KLV, Rows>), dim3(P), dim3(BL), 0, s, tours, n, rows_of, ct, k, \
                     max_iter, stats, cost, ws, stride)
    if (block == 1024) {
      TSPMI_CAND_BY_K(TSPMI_LAUNCH_CWS, 1024, k)
    } else if (block == 512) {
      TSPMI_CAND_BY_K(TSPMI_LAUNCH_CWS, 512, k)
    } else {
      TSPMI_CAND_BY_K(TSPMI_LAUNCH_CWS, 256, k)
    }
#undef TSPMI_LAUNCH_CWS
  }
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_large launch");
  return TSPMI_OK;
}

}  // namespace

extern "C" int tspmi_two_opt_large(int32_t* tours, int P, int n, const int32_t* D, int ld, const double* xy,
                                   const int32_t* table, int k, int max_iter, int64_t* stats, double* cost,
                                   int32_t* ws, void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_two_opt_large: negative shape P=%d n=%d", P, n);
  TSPMI_CHECK_ARG(table == nullptr || (k >= 1 && k <= 32), "tspmi_two_opt_large: k=%d outside [1, 32]", k);
  if (P == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(cost == nullptr || (n >= 4 && max_iter != 0), "tspmi_two_opt_large: costs need n >= 4 and max_iter != 0");
  if (n < 4 || max_iter == 0) return TSPMI_OK;  // reference: n < 4 or no budget is a copy
  TSPMI_CHECK_ARG(tours != nullptr && ws != nullptr, "tspmi_two_opt_large: null tours or workspace");
  TSPMI_CHECK_ARG((long)n * (long)n < (1L << 32), "tspmi_two_opt_large: n=%d too large for 32-bit scan indices", n);
  TSPMI_CHECK_ARG(table == nullptr || (long)(n - 2) * k < (1L << 31), "tspmi_two_opt_large: (n-2)*k too large");
  TSPMI_CHECK_ARG(table == nullptr || ((uintptr_t)table & 7) == 0, "tspmi_two_opt_large: table must be 8-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const int2* ct = reinterpret_cast<const int2*>(table);
  if (xy != nullptr) {
    TSPMI_CHECK_ARG(((uintptr_t)xy & 15) == 0, "tspmi_two_opt_large: xy must be 16-byte aligned");
    return launch_large(tours, P, n, CoordRows{reinterpret_cast<const double2*>(xy)}, ct, k, max_iter, stats, cost,
                        ws, s);
  }
  TSPMI_CHECK_ARG(D != nullptr && ld >= n, "tspmi_two_opt_large: null matrix or ld=%d < n=%d", ld, n);
  TSPMI_CHECK_ARG((long)n * ld < kMaxMatrixElems, "tspmi_two_opt_large: n * ld = %ld >= 2^30; pass coordinates",
                  (long)n * ld);
  return launch_large(tours, P, n, MatrixRows{D, ld}, ct, k, max_iter, stats, cost, ws, s);
}
