// Distance-matrix build, k-NN candidate tables and tour costs (gfx950).
#include "common.hpp"

namespace tspmi {

constexpr int kDistTile = 256;

__global__ __launch_bounds__(kDistTile) void dist_euc2d_kernel(const double* __restrict__ xy, int n,
                                                               int32_t* __restrict__ D, int ld,
                                                               int32_t* __restrict__ max_out) {
  __shared__ int s_max[kDistTile / kWave];
  const int u = blockIdx.y;
  const int v = blockIdx.x * kDistTile + threadIdx.x;
  int d = 0;
  if (v < n) {
    d = (u == v) ? 0 : euc2d_nint(xy[2 * u], xy[2 * u + 1], xy[2 * v], xy[2 * v + 1]);
    D[(size_t)u * ld + v] = d;
  }
  if (max_out != nullptr) {
    int m = d;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
      const int o = __shfl_xor(m, s, kWave);
      m = o > m ? o : m;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) s_max[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < kDistTile / kWave; ++w) m = s_max[w] > m ? s_max[w] : m;
      atomicMax(max_out, m);
    }
  }
}

// EUC_2D matrix from int32 coordinates by the integer distance form (euc2d_nint_i32); a pair
// whose per-axis difference exceeds kIntCoordSpan gets -1 and is counted in *bad_out.
__global__ __launch_bounds__(kDistTile) void dist_euc2d_i32_kernel(const int2* __restrict__ xy, int n,
                                                                   int32_t* __restrict__ D, int ld,
                                                                   int32_t* __restrict__ bad_out) {
  const int u = blockIdx.y;
  const int v = blockIdx.x * kDistTile + threadIdx.x;
  if (v >= n) return;
  const int2 a = xy[u], b = xy[v];
  const long dx = (long)a.x - b.x, dy = (long)a.y - b.y;
  const bool fits = dx >= -kIntCoordSpan && dx <= kIntCoordSpan && dy >= -kIntCoordSpan && dy <= kIntCoordSpan;
  D[(size_t)u * ld + v] = (u == v) ? 0 : (fits ? euc2d_nint_i32(a.x, a.y, b.x, b.y) : -1);
  if (!fits && bad_out != nullptr) atomicAdd(bad_out, 1);
}

// Other TSPLIB 2D metrics (src/distances/matrix.py:77-131, pairwise.py:62-102), each evaluated
// with the operations and order of the NumPy expression (no FMA contraction).
enum Metric : int { kEuc2d = 0, kCeil2d = 1, kMan2d = 2, kMax2d = 3, kAtt = 4, kGeo = 5 };

__device__ __forceinline__ double geo_radians(double c) {
#pragma clang fp contract(off)
  const double deg = __builtin_trunc(c);
  return 3.141592 * (deg + 5.0 * (c - deg) / 3.0) / 180.0;
}

template <int METRIC>
__device__ __forceinline__ int metric_distance(double xu, double yu, double xv, double yv) {
#pragma clang fp contract(off)
  if constexpr (METRIC == kEuc2d) {
    return euc2d_nint(xu, yu, xv, yv);
  } else if constexpr (METRIC == kCeil2d) {
    const double dx = xu - xv, dy = yu - yv;
    return (int)__builtin_ceil(__builtin_sqrt(dx * dx + dy * dy));
  } else if constexpr (METRIC == kMan2d) {
    return (int)(__builtin_fabs(xu - xv) + __builtin_fabs(yu - yv) + 0.5);
  } else if constexpr (METRIC == kMax2d) {
    return (int)(__builtin_fmax(__builtin_fabs(xu - xv), __builtin_fabs(yu - yv)) + 0.5);
  } else if constexpr (METRIC == kAtt) {
    const double dx = xu - xv, dy = yu - yv;
    const double r = __builtin_sqrt((dx * dx + dy * dy) / 10.0);
    const int t = (int)(r + 0.5);
    return (double)t < r ? t + 1 : t;
  } else {
    const double lat_i = geo_radians(xu), lon_i = geo_radians(yu);
    const double lat_j = geo_radians(xv), lon_j = geo_radians(yv);
    const double q1 = cos(lon_i - lon_j);
    const double q2 = cos(lat_i - lat_j);
    const double q3 = cos(lat_i + lat_j);
    double arg = 0.5 * ((1.0 + q1) * q2 - (1.0 - q1) * q3);
    arg = arg < -1.0 ? -1.0 : (arg > 1.0 ? 1.0 : arg);
    return (int)(6378.388 * acos(arg) + 1.0);
  }
}

template <int METRIC>
__global__ __launch_bounds__(kDistTile) void dist_metric_kernel(const double* __restrict__ xy, int n,
                                                                int32_t* __restrict__ D, int ld,
                                                                int32_t* __restrict__ max_out) {
  __shared__ int s_max[kDistTile / kWave];
  const int u = blockIdx.y;
  const int v = blockIdx.x * kDistTile + threadIdx.x;
  int d = 0;
  if (v < n) {
    d = (u == v) ? 0 : metric_distance<METRIC>(xy[2 * u], xy[2 * u + 1], xy[2 * v], xy[2 * v + 1]);
    D[(size_t)u * ld + v] = d;
  }
  if (max_out != nullptr) {
    int m = d;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
      const int o = __shfl_xor(m, s, kWave);
      m = o > m ? o : m;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) s_max[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < kDistTile / kWave; ++w) m = s_max[w] > m ? s_max[w] : m;
      atomicMax(max_out, m);
    }
  }
}

// One workgroup per row u: k rounds of a block argmin over 64-bit keys (D[u,v] << 32 | v),
// each round taking the smallest key above the previous one.  Keys are unique, so the result
// is the (distance, node index) order of src/distances/matrix.py:205-209.  A (k+1)-th round
// detects a row whose k-th and (k+1)-th distances tie: it is counted in ties_out and flagged
// in tie_flags[u] (both nullable).
constexpr int kKnnBlock = 256;

__global__ __launch_bounds__(kKnnBlock) void knn_kernel(const int32_t* __restrict__ D, int n, int ld, int k,
                                                        int32_t* __restrict__ out, int32_t* __restrict__ ties_out,
                                                        int32_t* __restrict__ tie_flags) {
  __shared__ uint64_t red[kKnnBlock / kWave];
  const int u = blockIdx.x;
  const int32_t* row = D + (size_t)u * ld;
  uint64_t prev = 0;
  bool have_prev = false;
  uint32_t kth_dist = 0;
  for (int r = 0; r <= k; ++r) {
    uint64_t best = kNoMove;
    for (int v = threadIdx.x; v < n; v += blockDim.x) {
      if (v == u) continue;
      const uint64_t key = ((uint64_t)(uint32_t)row[v] << 32) | (uint32_t)v;
      if ((!have_prev || key > prev) && key < best) best = key;
    }
    best = block_min_u64(best, red);
    if (r < k) {
      if (threadIdx.x == 0) out[(size_t)u * k + r] = (int32_t)(uint32_t)best;
      kth_dist = (uint32_t)(best >> 32);
    } else if (threadIdx.x == 0) {
      const bool tie = best != kNoMove && (uint32_t)(best >> 32) == kth_dist;
      if (tie && ties_out != nullptr) atomicAdd(ties_out, 1);
      if (tie_flags != nullptr) tie_flags[u] = tie ? 1 : 0;
    }
    prev = best;
    have_prev = true;
  }
}

// One wave per tour; exact int64 sum converted once to double.
constexpr int kCostBlock = 256;

__global__ __launch_bounds__(kCostBlock) void tour_cost_kernel(const int32_t* __restrict__ tours, int P, int n,
                                                               const int32_t* __restrict__ D, int ld, int closed,
                                                               double* __restrict__ out) {
  const int r = blockIdx.x * (kCostBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= P) return;
  const int32_t* t = tours + (size_t)r * n;
  int64_t s = 0;
  for (int x = lane; x + 1 < n; x += kWave) s += D[(size_t)clamp_id(t[x], ld) * ld + clamp_id(t[x + 1], ld)];
  if (closed && n > 1 && lane == 0) s += D[(size_t)clamp_id(t[n - 1], ld) * ld + clamp_id(t[0], ld)];
  s = wave_sum_i64(s);
  if (lane == 0) out[r] = (double)s;
}

}  // namespace tspmi

extern "C" int tspmi_dist_euc2d(const double* xy, int n, int32_t* D, int ld, int32_t* max_out, void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(n >= 0 && ld >= n && n <= 65535, "tspmi_dist_euc2d: bad shape n=%d ld=%d (n <= 65535)", n, ld);
  if (n == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(xy != nullptr && D != nullptr, "tspmi_dist_euc2d: null pointer");
  dim3 grid((n + kDistTile - 1) / kDistTile, n);
  hipLaunchKernelGGL(dist_euc2d_kernel, grid, dim3(kDistTile), 0, (hipStream_t)stream, xy, n, D, ld, max_out);
  TSPMI_CHECK_LAUNCH("tspmi_dist_euc2d launch");
  return TSPMI_OK;
}

extern "C" int tspmi_dist_euc2d_i32(const int32_t* xy, int n, int32_t* D, int ld, int32_t* bad_out, void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(n >= 0 && ld >= n && n <= 65535, "tspmi_dist_euc2d_i32: bad shape n=%d ld=%d (n <= 65535)", n, ld);
  if (n == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(xy != nullptr && D != nullptr, "tspmi_dist_euc2d_i32: null pointer");
  TSPMI_CHECK_ARG(((uintptr_t)xy & 7) == 0, "tspmi_dist_euc2d_i32: xy must be 8-byte aligned");
  dim3 grid((n + kDistTile - 1) / kDistTile, n);
  hipLaunchKernelGGL(dist_euc2d_i32_kernel, grid, dim3(kDistTile), 0, (hipStream_t)stream,
                     reinterpret_cast<const int2*>(xy), n, D, ld, bad_out);
  TSPMI_CHECK_LAUNCH("tspmi_dist_euc2d_i32 launch");
  return TSPMI_OK;
}

extern "C" int tspmi_knn(const int32_t* D, int n, int ld, int k, int32_t* out, int32_t* ties_out, void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(n >= 0 && ld >= n && k >= 0, "tspmi_knn: bad shape n=%d ld=%d k=%d", n, ld, k);
  const int k_eff = n > 1 ? (k < n - 1 ? k : n - 1) : 0;
  TSPMI_CHECK_ARG(k_eff == k, "tspmi_knn: k=%d must already be clamped to min(k, n-1)=%d", k, k_eff);
  if (n == 0 || k == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(D != nullptr && out != nullptr, "tspmi_knn: null pointer");
  hipLaunchKernelGGL(knn_kernel, dim3(n), dim3(kKnnBlock), 0, (hipStream_t)stream, D, n, ld, k, out, ties_out,
                     nullptr);
  TSPMI_CHECK_LAUNCH("tspmi_knn launch");
  return TSPMI_OK;
}

extern "C" int tspmi_knn_flagged(const int32_t* D, int n, int ld, int k, int32_t* out, int32_t* tie_flags,
                                 void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(n >= 0 && ld >= n && k >= 0, "tspmi_knn_flagged: bad shape n=%d ld=%d k=%d", n, ld, k);
  const int k_eff = n > 1 ? (k < n - 1 ? k : n - 1) : 0;
  TSPMI_CHECK_ARG(k_eff == k, "tspmi_knn_flagged: k=%d must already be clamped to min(k, n-1)=%d", k, k_eff);
  if (n == 0 || k == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(D != nullptr && out != nullptr && tie_flags != nullptr, "tspmi_knn_flagged: null pointer");
  hipLaunchKernelGGL(knn_kernel, dim3(n), dim3(kKnnBlock), 0, (hipStream_t)stream, D, n, ld, k, out, nullptr,
                     tie_flags);
  TSPMI_CHECK_LAUNCH("tspmi_knn_flagged launch");
  return TSPMI_OK;
}

extern "C" int tspmi_tour_cost(const int32_t* tours, int P, int n, const int32_t* D, int ld, int closed, double* out,
                               void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_tour_cost: negative shape");
  if (P == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(out != nullptr, "tspmi_tour_cost: null output");
  TSPMI_CHECK_ARG(n <= 1 || (tours != nullptr && D != nullptr && ld >= 1), "tspmi_tour_cost: null input");
  const int per_block = kCostBlock / kWave;
  hipLaunchKernelGGL(tour_cost_kernel, dim3((P + per_block - 1) / per_block), dim3(kCostBlock), 0,
                     (hipStream_t)stream, tours, P, n, D, ld, closed, out);
  TSPMI_CHECK_LAUNCH("tspmi_tour_cost launch");
  return TSPMI_OK;
}

extern "C" int tspmi_dist_metric(const double* xy, int n, int metric, int32_t* D, int ld, int32_t* max_out,
                                 void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(n >= 0 && ld >= n && n <= 65535, "tspmi_dist_metric: bad shape n=%d ld=%d (n <= 65535)", n, ld);
  TSPMI_CHECK_ARG(metric >= kEuc2d && metric <= kGeo, "tspmi_dist_metric: unknown metric %d", metric);
  if (n == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(xy && D, "tspmi_dist_metric: null pointer");
  const dim3 grid((n + kDistTile - 1) / kDistTile, n);
  hipStream_t s = (hipStream_t)stream;
  switch (metric) {
    case kEuc2d: hipLaunchKernelGGL(dist_metric_kernel<kEuc2d>, grid, dim3(kDistTile), 0, s, xy, n, D, ld, max_out); break;
    case kCeil2d: hipLaunchKernelGGL(dist_metric_kernel<kCeil2d>, grid, dim3(kDistTile), 0, s, xy, n, D, ld, max_out); break;
    case kMan2d: hipLaunchKernelGGL(dist_metric_kernel<kMan2d>, grid, dim3(kDistTile), 0, s, xy, n, D, ld, max_out); break;
    case kMax2d: hipLaunchKernelGGL(dist_metric_kernel<kMax2d>, grid, dim3(kDistTile), 0, s, xy, n, D, ld, max_out); break;
    case kAtt: hipLaunchKernelGGL(dist_metric_kernel<kAtt>, grid, dim3(kDistTile), 0, s, xy, n, D, ld, max_out); break;
    default: hipLaunchKernelGGL(dist_metric_kernel<kGeo>, grid, dim3(kDistTile), 0, s, xy, n, D, ld, max_out); break;
  }
  TSPMI_CHECK_LAUNCH("tspmi_dist_metric launch");
  return TSPMI_OK;
}
