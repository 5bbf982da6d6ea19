// Device building blocks of the GA operators, shared by the single-instance kernels (ga_ops.hip)
// and the multi-instance batch kernels (batch.hip).
#pragma once

#include "common.hpp"

namespace tspmi {

constexpr int kRowBlock = 256;
constexpr int kSortBlock = 1024;
constexpr int kSortMax = 4096;

// First minimum of fitness over one tournament row, np.argmin's rule (tournament_selection.py:125-127):
// the first NaN if any, else the first least value.
__device__ __forceinline__ int tournament_winner(const int32_t* __restrict__ cand_row, int k,
                                                 const double* __restrict__ fitness, int P) {
  int w = clamp_id(cand_row[0], P);
  uint64_t best = argmin_bits(fitness[w]);
  for (int c = 1; c < k; ++c) {
    const int x = clamp_id(cand_row[c], P);
    const uint64_t f = argmin_bits(fitness[x]);
    if (f < best) {
      best = f;
      w = x;
    }
  }
  return w;
}

// tournament_winner by one wave: lane c loads candidate c and its fitness, so the k candidate
// loads and the k fitness loads are each issued together instead of as k dependent pairs; the
// winner is the lowest-indexed candidate of least fitness (the sequential scan's first strict
// minimum; a NaN fitness wins as in np.argmin).  Every lane returns it.
__device__ __forceinline__ int wave_tournament(const int32_t* __restrict__ cand_row, int k,
                                               const double* __restrict__ fitness, int P) {
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t best = ~0ull;
  int w = 0;
  for (int c0 = 0; c0 < k; c0 += kWave) {
    const int c = c0 + lane;
    const int x = c < k ? clamp_id(cand_row[c], P) : 0;
    const uint64_t f = c < k ? argmin_bits(fitness[x]) : ~0ull;
    const uint64_t m = wave_min_u64(f);
    const uint64_t hit = __ballot(c < k && f == m);
    const int first = __builtin_ctzll(hit);
    const int xw = __shfl(x, first);
    if (c0 == 0 || m < best) {
      best = m;
      w = xw;
    }
  }
  return w;
}

// PMX child of (p1, p2) with cuts [c1, c2] into `child` (pmx_crossover.py:49-105, open
// permutation).  inv1/inv2 map node -> position in p1/p2.  Caller synchronises afterwards.
//
// The reference places each value p2[idx] of the segment that p1's segment lacks by walking
// pos -> inv2[p1[pos]] from idx until pos leaves [c1, c2].  Those walks are paths of the partial
// map S(pos) = inv2[p1[pos]] on segment positions: a start idx has no preimage inside the
// segment, so it never lies on a cycle.  With random parents a walk runs to ~150 dependent steps
// at n = 1000, and the longest one sets the workgroup's time.  When `jump` (c2 - c1 + 1 ints of
// LDS) is given, S is tabulated and squared in place, T <- T o T with exit positions fixed, until
// every start's entry has left the segment: ceil(log2(longest walk)) rounds of one barrier each,
// none at all when no value needs placing.  Each thread squares the entries of the start
// positions it owns, so it knows when its own starts are resolved.  An entry read while its
// owner rewrites it is either the old or the new value, both points of the same path no further
// than its exit, so the race never changes a result.  The placement -- the exit position of each
// walk -- is the reference's.
__device__ __forceinline__ void pmx_child(const int32_t* p1, const int32_t* p2, const int32_t* inv1, const int32_t* inv2, int n,
                          int c1, int c2, int32_t* child, int32_t* jump) {
  for (int x = threadIdx.x; x < n; x += blockDim.x) child[x] = (x >= c1 && x <= c2) ? p1[x] : -1;
  const int seg_len = c2 - c1 + 1;
  if (jump != nullptr && seg_len <= 32 * (int)blockDim.x) {
    uint32_t starts = 0;  // bit m: segment entry threadIdx.x + m * blockDim.x starts a walk
    bool pending = false;
    int m = 0;
    for (int e = threadIdx.x; e < seg_len; e += blockDim.x, ++m) {
      const int pv = inv1[p2[c1 + e]];
      const int t = inv2[p1[c1 + e]];
      jump[e] = t;
      if (pv < c1 || pv > c2) {
        starts |= 1u << m;
        pending |= (t >= c1 && t <= c2);
      }
    }
    pending = __syncthreads_or(pending);
    const int max_rounds = 32 - __clz(seg_len);  // 2^rounds >= seg_len >= any walk
    for (int r = 0; pending && r < max_rounds; ++r) {
      bool mine = false;
      m = 0;
      for (int e = threadIdx.x; e < seg_len; e += blockDim.x, ++m) {
        int t = jump[e];
        if (t >= c1 && t <= c2) {
          t = jump[t - c1];
          jump[e] = t;
        }
        mine |= ((starts >> m) & 1u) && t >= c1 && t <= c2;
      }
      pending = __syncthreads_or(mine);
    }
    m = 0;
    for (int e = threadIdx.x; e < seg_len; e += blockDim.x, ++m) {
      if (!((starts >> m) & 1u)) continue;
      const int pos = jump[e];
      if ((unsigned)pos < (unsigned)n && (pos < c1 || pos > c2)) child[pos] = p2[c1 + e];
    }
  } else {
    __syncthreads();
    for (int idx = c1 + threadIdx.x; idx <= c2; idx += blockDim.x) {
      const int val = p2[idx];
      const int pv = inv1[val];
      if (pv >= c1 && pv <= c2) continue;  // val already in the copied segment
      int pos = idx, steps = 0;
      while (pos >= c1 && pos <= c2 && steps <= seg_len) {
        pos = inv2[p1[pos]];
        ++steps;
      }
      if ((unsigned)pos < (unsigned)n && (pos < c1 || pos > c2)) child[pos] = val;
    }
  }
  __syncthreads();
  for (int x = threadIdx.x; x < n; x += blockDim.x)
    if (child[x] == -1) child[x] = p2[x];
}

__device__ __forceinline__ void load_parent(const int32_t* __restrict__ src, int n, int32_t* p, int32_t* inv) {
  for (int x = threadIdx.x; x < n; x += blockDim.x) {
    const int v = clamp_id(src[x], n);
    p[x] = v;
    inv[v] = x;
  }
}

// Both parents of a pair (rows a and b) into LDS with their inverses; four elements of each row per
// thread are loaded before any is stored, so a round waits on global memory once.
__device__ __forceinline__ void load_parents(const int32_t* __restrict__ a, const int32_t* __restrict__ b, int n,
                                             int32_t* p1, int32_t* inv1, int32_t* p2, int32_t* inv2) {
  constexpr int U = 4;
  for (int x0 = threadIdx.x; x0 < n; x0 += U * blockDim.x) {
    int va[U], vb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = x0 + u * blockDim.x;
      va[u] = x < n ? a[x] : 0;
      vb[u] = x < n ? b[x] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = x0 + u * blockDim.x;
      if (x < n) {
        const int v1 = clamp_id(va[u], n), v2 = clamp_id(vb[u], n);
        p1[x] = v1;
        inv1[v1] = x;
        p2[x] = v2;
        inv2[v2] = x;
      }
    }
  }
}

__device__ __forceinline__ void reverse_range(int32_t* t, int lo, int hi) {
  const int half = (hi - lo + 1) >> 1;
  for (int q = threadIdx.x; q < half; q += blockDim.x) {
    const int32_t tmp = t[lo + q];
    t[lo + q] = t[hi - q];
    t[hi - q] = tmp;
  }
}

template <class Dist>
__device__ __forceinline__ double block_cost(const int32_t* t, int n, const Dist& dist, int64_t* red) {
  // four edges per thread per round, their distance gathers issued together
  constexpr int U = 4;
  int64_t s = 0;
  if (n > 1) {
    for (int x0 = threadIdx.x; x0 < n; x0 += U * blockDim.x) {
      int d[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = x0 + u * blockDim.x;
        d[u] = 0;
        if (x < n) {  // a branch, not a select: a distance from coordinates is costly
          const int nx = (x + 1 == n) ? 0 : x + 1;
          d[u] = dist(t[x], t[nx]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) s += d[u];
    }
  }
  return (double)block_sum_i64(s, red);
}

// Inversion (when drawn), cost and store of child row r held in LDS `ch`.
template <class Dist>
__device__ __forceinline__ void finish_child(int r, int32_t* ch, int n, const Dist& dist,
                                             const double* __restrict__ probs, const int32_t* __restrict__ pos1,
                                             const int32_t* __restrict__ pos2, double rate,
                                             int32_t* __restrict__ offspring, double* __restrict__ off_fit,
                                             int64_t* red) {
  if (probs[r] < rate) {
    int lo = pos1[r], hi = pos2[r];
    if (lo > hi) {
      const int t = lo;
      lo = hi;
      hi = t;
    }
    lo = lo < 0 ? 0 : lo;
    hi = hi >= n ? n - 1 : hi;
    reverse_range(ch, lo, hi);
    __syncthreads();
  }
  const double cost = block_cost(ch, n, dist, red);
  for (int x = threadIdx.x; x < n; x += blockDim.x) offspring[(size_t)r * n + x] = ch[x];
  if (threadIdx.x == 0) off_fit[r] = cost;
}

// Fused tournament -> PMX -> inversion -> evaluate of child row r (pair q = r / 2, which shares
// its parents and cut with row r ^ 1) by one workgroup.  LDS: p1, p2, inv1, inv2, child, jump
// table (6n ints) in `lds`; `red` holds one int64 per wave and `s_w` two ints.  Both rows of a
// pair load both tournament winners; row 2q is PMX(p1, p2), row 2q + 1 is PMX(p2, p1), and an
// odd last row copies its own winner (pmx_crossover.py:128-171).  One workgroup per child halves
// the serial chain of a pair's workgroup (two PMX builds, costs and stores).
// Child costs come from `dist` (a MatrixDist or, for EUC_2D instances, a CoordDist: same values).
template <class Dist>
__device__ __forceinline__ void generation_child(int r, const int32_t* __restrict__ pop, const double* __restrict__ fitness, int P,
                                 int n, const Dist dist, const int32_t* __restrict__ tour_cand,
                                 int k, const int32_t* __restrict__ cut_lo, const int32_t* __restrict__ cut_hi,
                                 const double* __restrict__ probs, const int32_t* __restrict__ pos1,
                                 const int32_t* __restrict__ pos2, double rate, int32_t* __restrict__ offspring,
                                 double* __restrict__ off_fit, int32_t* lds, int64_t* red, int* s_w) {
  const int q = r >> 1;
  const int r0 = 2 * q;
  const bool single = (r0 + 1 >= P);
  int32_t* p1 = lds;
  int32_t* p2 = lds + n;
  int32_t* inv1 = lds + 2 * n;
  int32_t* inv2 = lds + 3 * n;
  int32_t* ch = lds + 4 * n;
  int32_t* jump = lds + 5 * n;

  const int wid = threadIdx.x / kWave;
  if (wid < 2 && (wid == 0 || !single)) {
    const int w = wave_tournament(tour_cand + (size_t)(r0 + wid) * k, k, fitness, P);
    if ((threadIdx.x & (kWave - 1)) == 0) s_w[wid] = w;
  }
  __syncthreads();
  if (single) {
    const int32_t* src = pop + (size_t)s_w[0] * n;
    for (int x = threadIdx.x; x < n; x += blockDim.x) ch[x] = clamp_id(src[x], n);
    __syncthreads();
    finish_child(r, ch, n, dist, probs, pos1, pos2, rate, offspring, off_fit, red);
    return;
  }
  load_parents(pop + (size_t)s_w[0] * n, pop + (size_t)s_w[1] * n, n, p1, inv1, p2, inv2);
  __syncthreads();
  int c1 = cut_lo[q], c2 = cut_hi[q];
  if (c1 > c2) {
    const int t = c1;
    c1 = c2;
    c2 = t;
  }
  c1 = c1 < 0 ? 0 : c1;
  c2 = c2 >= n ? n - 1 : c2;
  if (r == r0) pmx_child(p1, p2, inv1, inv2, n, c1, c2, ch, jump);
  else pmx_child(p2, p1, inv2, inv1, n, c1, c2, ch, jump);
  __syncthreads();
  finish_child(r, ch, n, dist, probs, pos1, pos2, rate, offspring, off_fit, red);
}


// Bitonic sort of (orderable bits, index) for M <= M2 <= kSortMax keys by one workgroup of kSortBlock
// threads; order[r] = index of rank r for r < keep.
__device__ __forceinline__ void sort_select_body(const double* __restrict__ keys, int M, int M2, int keep, int32_t* __restrict__ order,
                                 uint64_t* s_key, int32_t* s_idx) {
  for (int x = threadIdx.x; x < M2; x += blockDim.x) {
    s_key[x] = x < M ? orderable_bits(keys[x]) : ~0ull;
    s_idx[x] = x;
  }
  __syncthreads();
  for (int size = 2; size <= M2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int p = threadIdx.x; p < (M2 >> 1); p += blockDim.x) {
        const int i = 2 * stride * (p / stride) + (p % stride);
        const int j = i + stride;
        const bool up = (i & size) == 0;
        const uint64_t ki = s_key[i], kj = s_key[j];
        const int ii = s_idx[i], ij = s_idx[j];
        const bool greater = (ki > kj) || (ki == kj && ii > ij);
        if (greater == up) {
          s_key[i] = kj;
          s_key[j] = ki;
          s_idx[i] = ij;
          s_idx[j] = ii;
        }
      }
      __syncthreads();
    }
  }
  for (int x = threadIdx.x; x < keep; x += blockDim.x) order[x] = s_idx[x];
}


}  // namespace tspmi
