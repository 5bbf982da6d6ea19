// Population initialisers on device: SplitMix64 Fisher-Yates permutations and GRASP
// nearest-neighbour tours (gfx950).  One thread per row, exactly as the row recurrences of
// src/algorithms/strategies/initialization/random_init.py:40-55 and
// nearest_neighbor_init.py:31-99 are sequential; rows are independent.
#include "common.hpp"

namespace tspmi {

// SplitMix64 (Steele, Lea & Flood 2014): state += golden gamma, then two xor-shift-multiply
// mixing rounds; the bounded draw keeps the high 32 bits times the bound.
struct SplitMix64 {
  uint64_t s;
  __device__ __forceinline__ uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  __device__ __forceinline__ uint32_t bounded(uint32_t bound) {
    return (uint32_t)(((next() >> 32) * (uint64_t)bound) >> 32);
  }
};

constexpr int kInitBlock = 64;

__global__ __launch_bounds__(kInitBlock) void random_init_kernel(const uint64_t* __restrict__ seeds, int P,
                                                                 const int32_t* __restrict__ customers, int n,
                                                                 int32_t* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P) return;
  int32_t* row = out + (size_t)r * n;
  for (int x = 0; x < n; ++x) row[x] = customers[x];
  SplitMix64 g{seeds[r]};
  for (int i = n - 1; i > 0; --i) {
    const int j = (int)g.bounded((uint32_t)(i + 1));
    const int32_t t = row[i];
    row[i] = row[j];
    row[j] = t;
  }
}

// `work` holds, per row, unvisited[n] and pos_map[n] (node ids must lie in [0, n)).
__global__ __launch_bounds__(kInitBlock) void nn_init_kernel(const uint64_t* __restrict__ seeds, int P,
                                                             const int32_t* __restrict__ customers, int n,
                                                             const int32_t* __restrict__ cand, int k,
                                                             int32_t* __restrict__ work, int32_t* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P) return;
  int32_t* unvisited = work + (size_t)r * 2 * n;
  int32_t* pos_map = unvisited + n;
  int32_t* row = out + (size_t)r * n;
  for (int x = 0; x < n; ++x) pos_map[x] = -1;
  for (int x = 0; x < n; ++x) {
    const int c = clamp_id(customers[x], n);
    unvisited[x] = c;
    pos_map[c] = x;
  }
  SplitMix64 g{seeds[r]};
  int rem = n;
  int start = (int)g.bounded((uint32_t)rem);
  int current = unvisited[start];
  row[0] = current;
  int last = rem - 1;
  int last_node = unvisited[last];
  unvisited[start] = last_node;
  unvisited[last] = current;
  pos_map[last_node] = start;
  pos_map[current] = last;
  rem = last;
  // cand_order is a permutation of 0..k-1 rebuilt per step; k <= 64 keeps it in registers/scratch.
  int order[64];
  for (int step = 1; step < n; ++step) {
    int chosen = -1;
    if (k > 0) {
      for (int i = 0; i < k; ++i) order[i] = i;
      const int32_t* crow = cand + (size_t)current * k;
      for (int i = 0; i < k; ++i) {
        const int j = i + (int)g.bounded((uint32_t)(k - i));
        const int t = order[i];
        order[i] = order[j];
        order[j] = t;
        const int nb = crow[order[i]];
        if (nb >= 0 && nb < n) {
          const int p = pos_map[nb];
          if (p >= 0 && p < rem) {
            chosen = nb;
            break;
          }
        }
      }
    }
    if (chosen < 0) chosen = unvisited[(int)g.bounded((uint32_t)rem)];
    row[step] = chosen;
    const int p = pos_map[chosen];
    last = rem - 1;
    last_node = unvisited[last];
    unvisited[p] = last_node;
    unvisited[last] = chosen;
    pos_map[last_node] = p;
    pos_map[chosen] = last;
    rem = last;
    current = chosen;
  }
}


// Greedy nearest-neighbour tours (always the nearest unvisited city, lowest index on ties), one
// workgroup per tour: each step is a workgroup argmin of (D[current][j], j) over the unvisited j,
// with the visited set as an LDS bitmap of ceil(n / 32) words.
constexpr int kGreedyBlock = 256;

__global__ __launch_bounds__(kGreedyBlock) void nn_greedy_kernel(const int32_t* __restrict__ starts, int n,
                                                                 const int32_t* __restrict__ D, int ld,
                                                                 int32_t* __restrict__ out) {
  extern __shared__ uint32_t visited[];
  __shared__ uint64_t red[kGreedyBlock / kWave];
  const int words = (n + 31) >> 5;
  for (int w = threadIdx.x; w < words; w += kGreedyBlock) visited[w] = 0u;
  __syncthreads();
  int32_t* row = out + (size_t)blockIdx.x * n;
  int cur = clamp_id(starts[blockIdx.x], n);
  if (threadIdx.x == 0) {
    row[0] = cur;
    visited[cur >> 5] |= 1u << (cur & 31);
  }
  __syncthreads();
  for (int step = 1; step < n; ++step) {
    const int32_t* drow = D + (size_t)cur * ld;
    uint64_t best = ~0ull;
    for (int j = threadIdx.x; j < n; j += kGreedyBlock) {
      if (!((visited[j >> 5] >> (j & 31)) & 1u)) {
        const uint64_t key = ((uint64_t)(uint32_t)drow[j] << 32) | (uint32_t)j;
        best = key < best ? key : best;
      }
    }
    best = block_min_u64(best, red);
    cur = (int)(uint32_t)best;
    if (threadIdx.x == 0) {
      row[step] = cur;
      visited[cur >> 5] |= 1u << (cur & 31);
    }
    __syncthreads();
  }
}

}  // namespace tspmi

using namespace tspmi;

extern "C" int tspmi_random_init(const uint64_t* seeds, int P, const int32_t* customers, int n, int32_t* out,
                                 void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_random_init: negative shape");
  if (P == 0 || n == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(seeds && customers && out, "tspmi_random_init: null pointer");
  hipLaunchKernelGGL(random_init_kernel, dim3((P + kInitBlock - 1) / kInitBlock), dim3(kInitBlock), 0,
                     (hipStream_t)stream, seeds, P, customers, n, out);
  TSPMI_CHECK_LAUNCH("tspmi_random_init launch");
  return TSPMI_OK;
}

extern "C" int tspmi_nn_init(const uint64_t* seeds, int P, const int32_t* customers, int n, const int32_t* cand, int k,
                             int32_t* work, int32_t* out, void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0 && k >= 0 && k <= 64, "tspmi_nn_init: bad shape (k must be <= 64)");
  if (P == 0 || n == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(seeds && customers && work && out && (k == 0 || cand), "tspmi_nn_init: null pointer");
  hipLaunchKernelGGL(nn_init_kernel, dim3((P + kInitBlock - 1) / kInitBlock), dim3(kInitBlock), 0,
                     (hipStream_t)stream, seeds, P, customers, n, cand, k, work, out);
  TSPMI_CHECK_LAUNCH("tspmi_nn_init launch");
  return TSPMI_OK;
}

extern "C" int tspmi_nn_greedy_init(const int32_t* starts, int P, int n, const int32_t* D, int ld, int32_t* out,
                                    void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_nn_greedy_init: negative shape");
  if (P == 0 || n == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(starts && D && out && ld >= n, "tspmi_nn_greedy_init: null pointer or ld < n");
  const size_t lds = (size_t)((n + 31) / 32) * sizeof(uint32_t);
  TSPMI_CHECK_ARG(lds + 64 <= (size_t)kMaxLds, "tspmi_nn_greedy_init: n=%d exceeds the LDS visited bitmap", n);
  hipLaunchKernelGGL(nn_greedy_kernel, dim3(P), dim3(kGreedyBlock), lds, (hipStream_t)stream, starts, n, D, ld, out);
  TSPMI_CHECK_LAUNCH("tspmi_nn_greedy_init launch");
  return TSPMI_OK;
}
