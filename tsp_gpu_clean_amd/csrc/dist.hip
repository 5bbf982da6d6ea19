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

// One workgroup per row u: k rounds of a block argmin over 64-bit keys (D[u,v] << 32 | v),
// each round taking the smallest key above the previous one.  Keys are unique, so the result
// is the (distance, node index) order of src/distances/matrix.py:205-209.
constexpr int kKnnBlock = 256;

__global__ __launch_bounds__(kKnnBlock) void knn_kernel(const int32_t* __restrict__ D, int n, int ld, int k,
                                                        int32_t* __restrict__ out, int32_t* __restrict__ ties_out) {
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
    } else if (ties_out != nullptr && best != kNoMove && threadIdx.x == 0 && (uint32_t)(best >> 32) == kth_dist) {
      atomicAdd(ties_out, 1);
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

extern "C" int tspmi_knn(const int32_t* D, int n, int ld, int k, int32_t* out, int32_t* ties_out, void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(n >= 0 && ld >= n && k >= 0, "tspmi_knn: bad shape n=%d ld=%d k=%d", n, ld, k);
  const int k_eff = n > 1 ? (k < n - 1 ? k : n - 1) : 0;
  TSPMI_CHECK_ARG(k_eff == k, "tspmi_knn: k=%d must already be clamped to min(k, n-1)=%d", k, k_eff);
  if (n == 0 || k == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(D != nullptr && out != nullptr, "tspmi_knn: null pointer");
  hipLaunchKernelGGL(knn_kernel, dim3(n), dim3(kKnnBlock), 0, (hipStream_t)stream, D, n, ld, k, out, ties_out);
  TSPMI_CHECK_LAUNCH("tspmi_knn launch");
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
