// Device building blocks of the 2-opt kernels, shared by the single-instance kernels (two_opt.hip)
// and the multi-instance batch kernels (batch.hip).
#pragma once

#include "common.hpp"

namespace tspmi {

constexpr int kTwoOptBlock = 256;
constexpr int kTwoOptWaves = kTwoOptBlock / kWave;

struct TwoOptShared {
  uint64_t red[kTwoOptWaves];
  int64_t red64[kTwoOptWaves];
};

__device__ __forceinline__ int wrap_next(int x, int n) { return (x + 1 == n) ? 0 : x + 1; }

// Distance policies: a row-major int32 matrix, or EUC_2D coordinates with the distance computed
// on the fly by the same FP64 function that builds the matrix (bit-identical values).
struct MatrixDist {
  const int32_t* __restrict__ D;
  int ld;
  __device__ __forceinline__ int operator()(int u, int v) const { return D[(size_t)u * ld + v]; }
};

struct CoordDist {
  const double2* __restrict__ xy;
  __device__ __forceinline__ int operator()(int u, int v) const {
    const double2 a = xy[u], b = xy[v];
    return euc2d_nint(a.x, a.y, b.x, b.y);
  }
};

// Reverse s_tour[i+1 .. j] and refresh edge costs (and positions) of the touched range.
template <class Dist>
__device__ __forceinline__ void apply_move(int32_t* s_tour, int32_t* s_edge, int32_t* s_pos, int i, int j, int n,
                                           const Dist& dist) {
  const int half = (j - i) >> 1;
  for (int t = threadIdx.x; t < half; t += blockDim.x) {
    const int lo = i + 1 + t, hi = j - t;
    const int32_t tmp = s_tour[lo];
    s_tour[lo] = s_tour[hi];
    s_tour[hi] = tmp;
  }
  __syncthreads();
  for (int x = i + threadIdx.x; x <= j; x += blockDim.x) {
    const int a = s_tour[x];
    s_edge[x] = dist(a, s_tour[wrap_next(x, n)]);
    if (s_pos != nullptr && x > i) s_pos[a] = x;
  }
  __syncthreads();
}

__device__ __forceinline__ void apply_move(int32_t* s_tour, int32_t* s_edge, int32_t* s_pos, int i, int j, int n,
                                           const int32_t* __restrict__ D, int ld) {
  apply_move(s_tour, s_edge, s_pos, i, j, n, MatrixDist{D, ld});
}

__device__ __forceinline__ int row_at(const int32_t* __restrict__ row, int node) {
  return *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(row) + ((uint32_t)node << 2));
}

// Candidate-list steepest 2-opt on a packed candidate table.
//
// `ct` is (n, KP) int2 {neighbour, D[node, neighbour]} (neighbour -1 past the k real entries), so
// the first distance of every candidate pair comes with the neighbour id in one 8-byte load.  A
// wave covers 64 / KP consecutive rows per step (lane = row offset * KP + r), with no integer
// division; UNROLL steps are issued together so their loads overlap.  Within a lane the pairs
// are visited in increasing (i, r) scan order, so a strict `<` keeps the lane's first minimum and
// the packed (delta, i*k + r) key restores the reference tie-break across lanes.
template <int BLOCK, int KP, int UNROLL, class Dist>
__device__ void cand_two_opt_tour(int32_t* __restrict__ tour, int n, const Dist dist, const int2* __restrict__ ct,
                                  int k, int max_iter, int64_t* __restrict__ stats, int32_t* lds, uint64_t* red,
                                  int64_t* red64) {
  constexpr int NW = BLOCK / kWave;
  constexpr int RPW = kWave / KP;  // rows per wave step
  int32_t* s_tour = lds;           // n + 1 (s_tour[n] mirrors s_tour[0])
  int32_t* s_edge = lds + n + 1;   // n
  int32_t* s_pos = s_edge + n;     // n
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = tid / kWave;
  const int sub = lane / KP;       // row offset inside the wave step
  const int r = lane - sub * KP;   // candidate rank handled by this lane
  for (int x = tid; x < n; x += BLOCK) {
    const int v = clamp_id(tour[x], n);
    s_tour[x] = v;
    if (x == 0) s_tour[n] = v;
  }
  __syncthreads();
  for (int x = tid; x < n; x += BLOCK) {
    const int a = s_tour[x];
    s_edge[x] = dist(a, s_tour[x + 1]);
    s_pos[a] = x;
  }
  __syncthreads();

  const int rows = n - 2;
  int64_t my_deltas = 0;
  int moves = 0, passes = 0;
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    int bd = 0;
    uint32_t bf = 0xffffffffu;
    for (int base = wid * RPW * UNROLL; base < rows; base += NW * RPW * UNROLL) {
      int2 c[UNROLL];
      int row[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        row[u] = base + u * RPW + sub;
        const int a = s_tour[row[u] < rows ? row[u] : 0];
        c[u] = ct[(size_t)a * KP + r];
      }
      int g[UNROLL], jj[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const int i = row[u];
        const int nb = c[u].x;
        int j = (nb >= 0 && nb < n) ? s_pos[nb] : -1;
        const int lim = (i == 0) ? n - 1 : n;
        j = (r < k && i < rows && j > i + 1 && j < lim) ? j : -1;
        jj[u] = j;
        const int b = s_tour[i < rows ? i + 1 : 0];
        g[u] = dist(b, s_tour[j >= 0 ? j + 1 : 0]);
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const int j = jj[u];
        if (j >= 0) {
          const int i = row[u];
          const int delta = c[u].y + g[u] - s_edge[i] - s_edge[j];
          ++my_deltas;
          if (delta < bd) {
            bd = delta;
            bf = (uint32_t)i * (uint32_t)k + (uint32_t)r;
          }
        }
      }
    }
    uint64_t best = bd < 0 ? pack_key(bd, bf) : kNoMove;
    best = block_min_u64(best, red);
    if (best == kNoMove) break;
    const uint32_t idx = key_index(best);
    const int i = (int)(idx / (uint32_t)k);
    const int j = s_pos[ct[(size_t)s_tour[i] * KP + (int)(idx - (uint32_t)i * (uint32_t)k)].x];
    __syncthreads();
    apply_move(s_tour, s_edge, s_pos, i, j, n, dist);
    ++moves;
  }
  for (int x = tid; x < n; x += BLOCK) tour[x] = s_tour[x];
  if (stats != nullptr) {
    const int64_t total = block_sum_i64(my_deltas, red64);
    if (tid == 0) {
      atomicAdd((unsigned long long*)&stats[0], (unsigned long long)total);
      atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
      atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
    }
  }
}




}  // namespace tspmi
