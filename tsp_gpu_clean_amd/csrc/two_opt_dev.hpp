// Device building blocks of the 2-opt kernels, shared by the single-instance kernels (two_opt.hip)
// and the multi-instance batch kernels (batch.hip).
//
// The per-tour bodies are always inlined: once a body has several callers in one source file the
// compiler may emit it as an out-of-line call, and the call ABI then costs the kernel its
// registers and occupancy (the batch elite kernel went from 87 VGPRs and 5 waves per SIMD to 152
// and 3 that way).
#pragma once

#include "common.hpp"

namespace tspmi {

constexpr int kTwoOptBlock = 256;
constexpr int kTwoOptWaves = kTwoOptBlock / kWave;

struct TwoOptShared {
  uint64_t red[2 * kTwoOptWaves];  // the full-neighbourhood scan uses both halves
  int64_t red64[kTwoOptWaves];
};

__device__ __forceinline__ int wrap_next(int x, int n) { return (x + 1 == n) ? 0 : x + 1; }

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

// Tour (n node ids) into LDS, s_tour[n] mirroring s_tour[0]; four ids per thread are loaded before
// any is stored, so a round waits on global memory once.
template <int BLOCK>
__device__ __forceinline__ void load_tour_lds(const int32_t* __restrict__ tour, int n, int32_t* s_tour) {
  constexpr int U = 4;
  for (int x0 = (int)threadIdx.x; x0 < n; x0 += U * BLOCK) {
    int v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = x0 + u * BLOCK;
      v[u] = x < n ? tour[x] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = x0 + u * BLOCK;
      if (x < n) {
        const int c = clamp_id(v[u], n);
        s_tour[x] = c;
        if (x == 0) s_tour[n] = c;
      }
    }
  }
}

// int32 words of LDS that cand_two_opt_tour uses (tour, edge costs, positions), rounded up to an
// even count so that int2 coordinates can follow.
__host__ __device__ constexpr int cand_lds_words(int n) { return (3 * n + 2) & ~1; }

// Copies the n int32 coordinate pairs of an instance into LDS after cand_two_opt_tour's arrays and
// returns the distance policy that reads them there (no texture-addresser cycles for the scan's
// coordinate reads).  The copies are visible to the scan after the barrier that follows its tour
// load; the caller must have finished with that LDS (a barrier after the previous use).
template <int BLOCK>
__device__ __forceinline__ IntCoordDist stage_int_coords(const int2* __restrict__ xy, int n, int32_t* lds) {
  int2* s_xy = reinterpret_cast<int2*>(lds + cand_lds_words(n));
  for (int x = (int)threadIdx.x; x < n; x += BLOCK) s_xy[x] = xy[x];
  return IntCoordDist{s_xy};
}

// Lanes per table row of cand_two_opt_tour for k candidates: the smallest instantiated width >= k.
inline int cand_lane_group(int k) { return k <= 8 ? 8 : k <= 10 ? 10 : k <= 12 ? 12 : k <= 16 ? 16 : k <= 20 ? 20 : 32; }

// Expands LAUNCH(BL, KP, KL) for the lane group of runtime `k` (KP = the packed table's row width).
#define TSPMI_CAND_BY_K(LAUNCH, BL, k)    \
  switch (cand_lane_group(k)) {           \
    case 8: LAUNCH(BL, 8, 8); break;      \
    case 10: LAUNCH(BL, 16, 10); break;   \
    case 12: LAUNCH(BL, 16, 12); break;   \
    case 16: LAUNCH(BL, 16, 16); break;   \
    case 20: LAUNCH(BL, 32, 20); break;   \
    default: LAUNCH(BL, 32, 32); break;   \
  }

// Candidate-list steepest 2-opt on a packed candidate table.
//
// `ct` is (n, KP) int2 {neighbour, D[node, neighbour]} (neighbour -1 past the k real entries), so
// the first distance of every candidate pair comes with the neighbour id in one 8-byte load.  A
// wave covers RPW = 64 / KL consecutive rows per step (lane = row offset * KL + r, KL <= KP lanes
// per row, KL = k where an instance exists so that no lane idles on the table's padding; the
// 64 - RPW * KL spare lanes sit out), with the division by the constant KL done at compile time;
// UNROLL steps are issued together so their loads overlap.  Within a lane the pairs are visited
// in increasing (i, r) scan order, so a strict `<` keeps the lane's first minimum and the packed
// (delta, i*k + r) key restores the reference tie-break across lanes.
//
// `cost` (nullable) receives the closed cost of the final tour, the sum of its edge costs.
// `src` (nullable) is where the start tour is read from; the refined tour is always written to
// `tour` (so a survivor can be copied and refined in one step).
// `pass_log` (nullable) receives {tour cost, wall clock} int64 pairs: entry 0 at the start, entry
// m after the m-th move (m <= log_moves; later moves are not logged), the cost kept exactly as the
// start cost plus the applied deltas and the clock read with wall_clock64() (the device's
// constant-rate counter, the same on every CU).
template <int BLOCK, int KP, int UNROLL, int KL = KP, class Dist>
__device__ __forceinline__ void cand_two_opt_tour(int32_t* __restrict__ tour, int n, const Dist dist, const int2* __restrict__ ct,
                                  int k, int max_iter, int64_t* __restrict__ stats, int32_t* lds, uint64_t* red,
                                  int64_t* red64, double* __restrict__ cost = nullptr,
                                  int64_t* __restrict__ pass_log = nullptr, int log_moves = 0,
                                  const int32_t* __restrict__ src = nullptr) {
  static_assert(KL <= KP && KL <= kWave, "lanes per row exceed the table row");
  constexpr int NW = BLOCK / kWave;
  constexpr int RPW = kWave / KL;  // rows per wave step
  int32_t* s_tour = lds;           // n + 1 (s_tour[n] mirrors s_tour[0])
  int32_t* s_edge = lds + n + 1;   // n
  int32_t* s_pos = s_edge + n;     // n
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = wave_id();
  const int sub = lane / KL;       // row offset inside the wave step (RPW for a spare lane)
  const int r = lane - sub * KL;   // candidate rank handled by this lane
  load_tour_lds<BLOCK>(src != nullptr ? src : tour, n, s_tour);
  __syncthreads();
  {
    // four positions per thread per round, their distance gathers issued together
    constexpr int U = 4;
    for (int x0 = tid; x0 < n; x0 += U * BLOCK) {
      int a[U], d[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = x0 + u * BLOCK;
        a[u] = d[u] = 0;
        if (x < n) {  // a branch, not a select: a distance from coordinates is costly
          a[u] = s_tour[x];
          d[u] = dist(a[u], s_tour[x + 1]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = x0 + u * BLOCK;
        if (x < n) {
          s_edge[x] = d[u];
          s_pos[a[u]] = x;
        }
      }
    }
  }
  __syncthreads();

  const int rows = n - 2;
  int64_t my_deltas = 0;
  int moves = 0, passes = 0;
  int64_t run_cost = 0;
  if (pass_log != nullptr) {
    int64_t part = 0;
    for (int x = tid; x < n; x += BLOCK) part += s_edge[x];
    run_cost = block_sum_i64(part, red64);
    if (tid == 0) {
      pass_log[0] = run_cost;
      pass_log[1] = (int64_t)wall_clock64();
    }
  }
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    int bd = 0;
    uint32_t bf = 0xffffffffu;
    for (int base = wid * RPW * UNROLL; base < rows; base += NW * RPW * UNROLL) {
      int2 c[UNROLL];
      int row[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        row[u] = sub < RPW ? base + u * RPW + sub : rows;
        const int a = s_tour[row[u] < rows ? row[u] : 0];
        // 32-bit byte offset from the table base (n * KP * 8 < 2^32 for every n served here)
        c[u] = *reinterpret_cast<const int2*>(reinterpret_cast<const char*>(ct) +
                                              ((((uint32_t)a * KP) + (uint32_t)r) << 3));
      }
      int g[UNROLL], jj[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const int i = row[u];
        const int nb = c[u].x;
        const bool nbv = (unsigned)nb < (unsigned)n;
        int j = s_pos[nbv ? nb : 0];  // read unconditionally (no branch), discarded when nb is padding
        const int lim = (i == 0) ? n - 1 : n;
        j = (nbv && r < k && i < rows && j > i + 1 && j < lim) ? j : -1;
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
    if (pass_log != nullptr) {
      run_cost += (int32_t)((uint32_t)(best >> 32) ^ 0x80000000u);  // the applied delta
      if (tid == 0 && moves <= log_moves) {
        pass_log[2 * moves] = run_cost;
        pass_log[2 * moves + 1] = (int64_t)wall_clock64();
      }
    }
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
  if (cost != nullptr) {
    int64_t part = 0;
    for (int x = tid; x < n; x += BLOCK) part += s_edge[x];
    const int64_t total = block_sum_i64(part, red64);
    if (tid == 0) *cost = (double)total;
  }
}

// int32 words of LDS that rec_two_opt_tour uses for an n-city tour.
__host__ __device__ constexpr int rec_lds_words(int n) { return 7 * n + 1; }

// Candidate-list steepest 2-opt for EUC_2D instances with int32 coordinates whose per-axis span is
// at most kIntCoordSpan, with node-indexed records in LDS.
//
// Same pairs, deltas, tie-break, moves, counters and outputs as cand_two_opt_tour with
// IntCoordDist.  LDS holds, per node u, the record {position of u, cost of the edge from u to its
// successor v, coordinates of v} (16 bytes), the coordinates by node and the tour.  A lane takes
// node a and rank r (lane = node offset * KL + r, as there), loads the table entry {nb, D[a][nb]}
// and the record of a together, then the record of nb: position i of a, position j of nb, the
// two removed edges and the coordinates of the two new edge's endpoints t_{i+1} and t_{j+1} are
// all in hand, so a pair costs one global load and two LDS reads, two levels deep, with no
// position-to-node lookups.  The evaluation is branch-free and a lane keeps the smallest whole key
// (delta, i * k + r): the nodes are not visited in scan order, and the key restores the
// reference's order.  A move reverses the tour between i+1 and j and rewrites the records of the
// nodes at positions i .. j from the coordinates' copy in LDS.
// LDS: rec_lds_words(n) int32 words, 16-byte aligned.
template <int BLOCK, int KP, int UNROLL, int KL = KP>
__device__ __forceinline__ void rec_two_opt_tour(int32_t* __restrict__ tour, int n, const int2* __restrict__ xy,
                                                 const int2* __restrict__ ct, int k, int max_iter,
                                                 int64_t* __restrict__ stats, int32_t* lds, uint64_t* red,
                                                 int64_t* red64, double* __restrict__ cost = nullptr,
                                                 int64_t* __restrict__ pass_log = nullptr, int log_moves = 0,
                                                 const int32_t* __restrict__ src = nullptr) {
  static_assert(KL <= KP && KL <= kWave, "lanes per row exceed the table row");
  constexpr int NW = BLOCK / kWave;
  constexpr int RPW = kWave / KL;  // nodes per wave step
  int4* s_rec = reinterpret_cast<int4*>(lds);  // n
  const int2* s_xy = reinterpret_cast<const int2*>(lds + 4 * n);  // n
  int32_t* s_tour = lds + 6 * n;                                    // n + 1 (s_tour[n] mirrors s_tour[0])
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = wave_id();
  const int sub = lane / KL;      // node offset inside the wave step (RPW for a spare lane)
  const int r = lane - sub * KL;  // candidate rank handled by this lane
  load_tour_lds<BLOCK>(src != nullptr ? src : tour, n, s_tour);
  for (int x = tid; x < n; x += BLOCK) reinterpret_cast<int2*>(lds + 4 * n)[x] = xy[x];
  __syncthreads();
  for (int x = tid; x < n; x += BLOCK) {
    const int u = s_tour[x], v = s_tour[x + 1];
    const int2 pu = s_xy[u], pv = s_xy[v];
    s_rec[u] = make_int4(x, euc2d_nint_i32(pu.x, pu.y, pv.x, pv.y), pv.x, pv.y);
  }
  __syncthreads();

  const int rows = n - 2;
  int deltas = 0;  // admissible pairs evaluated by this lane
  int moves = 0, passes = 0;
  int64_t run_cost = 0;
  if (pass_log != nullptr) {
    int64_t part = 0;
    for (int x = tid; x < n; x += BLOCK) part += s_rec[x].y;
    run_cost = block_sum_i64(part, red64);
    if (tid == 0) {
      pass_log[0] = run_cost;
      pass_log[1] = (int64_t)wall_clock64();
    }
  }
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    uint64_t best = kNoMove;
    for (int base = wid * RPW * UNROLL; base < n; base += NW * RPW * UNROLL) {
      int2 c[UNROLL];
      int4 ra[UNROLL];
      bool va[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const int a = base + u * RPW + sub;
        va[u] = sub < RPW && a < n;
        const int ac = va[u] ? a : 0;
        // 32-bit byte offset from the table base (n * KP * 8 < 2^32 for every n served here)
        c[u] = *reinterpret_cast<const int2*>(reinterpret_cast<const char*>(ct) +
                                              ((((uint32_t)ac * KP) + (uint32_t)r) << 3));
        ra[u] = s_rec[ac];
      }
      int4 rb[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const int nb = c[u].x;
        va[u] = va[u] && r < k && (unsigned)nb < (unsigned)n;
        rb[u] = s_rec[va[u] ? nb : 0];
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const int i = ra[u].x, j = rb[u].x;
        const int lim = (i == 0) ? n - 1 : n;
        const bool ok = va[u] && i < rows && j > i + 1 && j < lim;
        const int g = euc2d_nint_i32(ra[u].z, ra[u].w, rb[u].z, rb[u].w);
        const int delta = c[u].y + g - ra[u].y - rb[u].y;
        deltas += ok ? 1 : 0;
        const uint64_t key = (ok && delta < 0) ? pack_key(delta, (uint32_t)i * (uint32_t)k + (uint32_t)r) : kNoMove;
        best = key < best ? key : best;
      }
    }
    best = block_min_u64(best, red);
    if (best == kNoMove) break;
    const uint32_t idx = key_index(best);
    const int i = (int)(idx / (uint32_t)k);
    const int j = s_rec[ct[(size_t)s_tour[i] * KP + (int)(idx - (uint32_t)i * (uint32_t)k)].x].x;
    __syncthreads();  // every thread has read i and j before the tour changes
    const int half = (j - i) >> 1;
    for (int t = tid; t < half; t += BLOCK) {
      const int lo = i + 1 + t, hi = j - t;
      const int32_t tmp = s_tour[lo];
      s_tour[lo] = s_tour[hi];
      s_tour[hi] = tmp;
    }
    __syncthreads();
    for (int x = i + tid; x <= j; x += BLOCK) {
      const int u = s_tour[x], v = s_tour[x + 1];
      const int2 pu = s_xy[u], pv = s_xy[v];
      s_rec[u] = make_int4(x, euc2d_nint_i32(pu.x, pu.y, pv.x, pv.y), pv.x, pv.y);
    }
    __syncthreads();
    ++moves;
    if (pass_log != nullptr) {
      run_cost += (int32_t)((uint32_t)(best >> 32) ^ 0x80000000u);  // the applied delta
      if (tid == 0 && moves <= log_moves) {
        pass_log[2 * moves] = run_cost;
        pass_log[2 * moves + 1] = (int64_t)wall_clock64();
      }
    }
  }
  for (int x = tid; x < n; x += BLOCK) tour[x] = s_tour[x];
  if (stats != nullptr) {
    const int64_t total = block_sum_i64(deltas, red64);
    if (tid == 0) {
      atomicAdd((unsigned long long*)&stats[0], (unsigned long long)total);
      atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
      atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
    }
  }
  if (cost != nullptr) {
    int64_t part = 0;
    for (int x = tid; x < n; x += BLOCK) part += s_rec[x].y;
    const int64_t total = block_sum_i64(part, red64);
    if (tid == 0) *cost = (double)total;
  }
}

// Node-indexed records {position, successor, D[node][successor], 0} of the tour in s_tour, four
// positions per thread per round with their distance gathers issued together.
template <int BLOCK, class Dist>
__device__ __forceinline__ void build_node_records(int4* s_nd, const int32_t* s_tour, int n, const Dist& dist) {
  constexpr int U = 4;
  for (int x0 = (int)threadIdx.x; x0 < n; x0 += U * BLOCK) {
    int a[U], b[U], d[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = x0 + u * BLOCK;
      a[u] = b[u] = d[u] = 0;
      if (x < n) {
        a[u] = s_tour[x];
        b[u] = s_tour[x + 1];
        d[u] = dist(a[u], b[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int x = x0 + u * BLOCK;
      if (x < n) s_nd[a[u]] = make_int4(x, b[u], d[u], 0);
    }
  }
}

// Edge costs of node-indexed records {position, successor, D[node][successor], 0} that a move left
// stale: the edges at positions lo .. hi (none when lo > hi).  Called between a pass's distance
// gathers and its first use of the costs, so the two sets of gathers are in flight together and a
// pass waits on global memory once.
template <int BLOCK, class Dist>
__device__ __forceinline__ void refresh_edge_costs(int4* s_nd, const int32_t* s_tour, int lo, int hi,
                                                   const Dist& dist) {
  constexpr int ZU = 4;
  for (int x0 = lo + (int)threadIdx.x; x0 <= hi; x0 += ZU * BLOCK) {
    int u[ZU], z[ZU];
#pragma unroll
    for (int q = 0; q < ZU; ++q) {
      const int x = x0 + q * BLOCK;
      const bool in = x <= hi;
      u[q] = s_tour[in ? x : 0];
      z[q] = dist(u[q], s_tour[in ? x + 1 : 1]);
      u[q] = in ? u[q] : -1;
    }
#pragma unroll
    for (int q = 0; q < ZU; ++q)
      if (u[q] >= 0) s_nd[u[q]].z = z[q];
  }
}

// Candidate-list steepest 2-opt for a few tours at a time (the GA's elite set), with the pairs in
// registers and the tour in node-indexed LDS.
//
// Same pairs, deltas, tie-break and move as cand_two_opt_tour.  With one tour per workgroup and
// pass after pass on it, that kernel's scan is a chain of two dependent global loads (candidate
// row, then D[b][d]) per UNROLL rows and sets the pass time.  Here the (node, candidate) pairs
// never change, so each thread loads its PP pairs once: pair p = tid + m * BLOCK is node a = p / k,
// rank r = p % k, kept as (a | c << 16) and D[a][c].  LDS holds s_tour (position -> node, n + 1
// entries) and one int4 per node {position, successor, D[node][successor], 0}, so a pair costs two
// 16-byte LDS reads and one gather D[b][d] per pass, all issued together.  Pairs are not visited
// in scan order, so a thread keeps the minimum of the whole key (delta, i * k + r); the move's j
// rides in the key's low 13 bits under the scan index (n <= 8192, n * k <= 2^19), which leaves the
// order unchanged and gives every thread the move straight from the reduction.
//
// LDS: 20 n + 4 bytes.  Requires n * k <= PP * BLOCK; at BLOCK = 1024, PP = 20 is the most that
// stays within 128 VGPRs without spilling.
constexpr int kRegJBits = 13;

template <int BLOCK, int PP, class Dist>
__device__ __forceinline__ void reg_two_opt_tour(int32_t* __restrict__ tour, int n, const Dist dist, const int2* __restrict__ ct,
                                 int kp, int k, int max_iter, int64_t* __restrict__ stats, double* __restrict__ cost,
                                 int32_t* lds, uint64_t* red, int64_t* red64) {
  int4* s_nd = reinterpret_cast<int4*>(lds);  // n
  int32_t* s_tour = lds + 4 * n;              // n + 1 (s_tour[n] mirrors s_tour[0])
  const int tid = threadIdx.x;
  load_tour_lds<BLOCK>(tour, n, s_tour);
  uint32_t ac[PP];
  int dac[PP];
  const int total = n * k;
  // (a, r) of pair tid + m * BLOCK by stepping quotient and remainder (one division, not PP: a
  // runtime integer division is a long instruction sequence); all PP table loads are issued
  // before any is tested (a branch per load would serialise them).
  int2 e[PP];
  int pa[PP];
  {
    const int qb = BLOCK / k, rb = BLOCK - qb * k;
    int a = tid / k, r = tid - a * k;
#pragma unroll
    for (int m = 0; m < PP; ++m) {
      pa[m] = a;
      const int p = tid + m * BLOCK;
      e[m] = ct[(size_t)(p < total ? a : 0) * kp + (p < total ? r : 0)];
      a += qb;
      r += rb;
      if (r >= k) {
        r -= k;
        ++a;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < PP; ++m) {
    const int p = tid + m * BLOCK;
    const bool ok = p < total && e[m].x >= 0 && e[m].x < n;
    ac[m] = ok ? ((uint32_t)pa[m] | ((uint32_t)e[m].x << 16)) : 0xffffffffu;
    dac[m] = ok ? e[m].y : 0;
  }
  __syncthreads();
  build_node_records<BLOCK>(s_nd, s_tour, n, dist);
  __syncthreads();

  const int rows = n - 2;
  int64_t my_deltas = 0;
  int moves = 0, passes = 0;
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    uint64_t best = kNoMove;
#pragma unroll
    for (int m = 0; m < PP; ++m) {
      const uint32_t w = ac[m];
      const int a = (int)(w & 0xffffu), c = (int)(w >> 16);
      const int4 na = s_nd[w == 0xffffffffu ? 0 : a];
      const int4 nc = s_nd[w == 0xffffffffu ? 0 : c];
      const int i = na.x, j = nc.x;
      const int lim = (i == 0) ? n - 1 : n;
      const bool ok = (w != 0xffffffffu) & ((unsigned)i < (unsigned)rows) & (j > i + 1) & (j < lim);
      const int delta = dac[m] + dist(na.y, nc.y) - na.z - nc.z;
      my_deltas += ok;
      if (ok & (delta < 0)) {
        const uint32_t r = (uint32_t)(tid + m * BLOCK - a * k);
        const uint64_t key = pack_key(delta, (((uint32_t)i * (uint32_t)k + r) << kRegJBits) | (uint32_t)j);
        best = key < best ? key : best;
      }
    }
    best = block_min_u64(best, red);
    if (best == kNoMove) break;
    const uint32_t idx = key_index(best);
    const int i = (int)((idx >> kRegJBits) / (uint32_t)k);
    const int j = (int)(idx & ((1u << kRegJBits) - 1u));
    // Reverse s_tour[i+1 .. j] (s_tour[0] and its mirror never move), then refresh the records of
    // the nodes now at positions i .. j.
    const int half = (j - i) >> 1;
    for (int t = tid; t < half; t += BLOCK) {
      const int lo = i + 1 + t, hi = j - t;
      const int32_t tmp = s_tour[lo];
      s_tour[lo] = s_tour[hi];
      s_tour[hi] = tmp;
    }
    __syncthreads();
    for (int x = i + tid; x <= j; x += BLOCK) {
      const int u = s_tour[x], v = s_tour[x + 1];
      s_nd[u] = make_int4(x, v, dist(u, v), 0);
    }
    __syncthreads();
    ++moves;
  }
  for (int x = tid; x < n; x += BLOCK) tour[x] = s_tour[x];
  if (stats != nullptr) {
    const int64_t tot = block_sum_i64(my_deltas, red64);
    if (tid == 0) {
      atomicAdd((unsigned long long*)&stats[0], (unsigned long long)tot);
      atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
      atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
    }
  }
  if (cost != nullptr) {
    int64_t part = 0;
    for (int x = tid; x < n; x += BLOCK) part += s_nd[x].z;
    const int64_t tot = block_sum_i64(part, red64);
    if (tid == 0) *cost = (double)tot;
  }
}




// Candidate-list steepest 2-opt of one tour by a cluster of G workgroups.
//
// Same pairs, deltas, tie-break and moves as reg_two_opt_tour.  With one workgroup per tour a
// pass runs on one CU, which is issue- and gather-bound (20000 pairs at n = 1000, k = 20) while
// the other CUs idle; here G workgroups share the tour's pairs (pair p = g * BLOCK + tid + m * G *
// BLOCK, kept in registers as before) and each holds its own copy of the tour in LDS.  Per pass
// every workgroup publishes its best key as one 64-bit word in its own line of the tour's
// workspace; wave 0 of every workgroup polls the G words, one lane each, until all carry this
// pass's tag, takes their minimum, and every workgroup applies that move to its own copy.
// Nothing is contended (each word has one writer, on its own 128-byte line) and the data is its
// own flag: a real key has bit 63 clear (its delta is negative), so bit 63 holds a tag, and "no
// move" is kClusterNone.  Pass t uses slot t & 1 with tag ((t >> 1) + 1) & 1, so consecutive uses
// of a slot alternate tags and the zeroed workspace (tag 0) never reads as pass 0 or 1.  A slot
// is rewritten two passes later, by which time every workgroup has read it: a workgroup only
// reaches pass t after every other one has published pass t - 1, which each does after reading
// pass t - 2.  The wait is bounded: a workgroup that spins kClusterSpinLimit times (its cluster
// was not co-resident) raises the tour's error word, the others see it and leave, and the tour
// is not written; the caller reads the word.  The word is checked on the first poll of every
// wait too, so a word left set (or set by the caller) makes the cluster leave at once.  The last workgroup to leave (a departure counter)
// clears the slots and the counter for the next launch.
//
// Workspace per tour (int64 words, lines of 16): line g < G = slots of workgroup g (2 words),
// line 2 * kClusterMaxG = {error, departures}.
constexpr int kClusterMaxG = 16;
constexpr int kClusterLine = 16;
constexpr int kClusterWords = (2 * kClusterMaxG + 1) * kClusterLine;
constexpr unsigned kClusterSpinLimit = 1u << 22;
constexpr uint64_t kClusterFail = ~1ull;           // never a key: a key's delta is negative
constexpr uint64_t kClusterNone = ~0ull >> 1;      // above every key, bit 63 clear
constexpr uint64_t kClusterTag = 1ull << 63;

__device__ __forceinline__ uint64_t load_relaxed(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

//
// `best` (nullable) receives the index of the first minimum of the P tours' costs (cost_all,
// written with agent-scope stores): the last tour to finish, counted in word 2 of tour 0's
// control line (ws0), takes the minimum and clears the count.
//
// With int32 coordinates (Dist = IntCoordDist, each axis spanning at most kIntCoordSpan), the
// workgroup copies them into LDS and a node record is {position, cost of the edge to the
// successor, the successor's coordinates}: a pair's two record reads then give positions, both
// removed edges and the new edge's endpoints, the new distance is computed, and a pass touches no
// global memory but the cluster's key words (with the matrix, {position, successor, edge cost}
// and a gather D[b][d] per pair, the stale edge costs refreshed in flight with those gathers).
// LDS: 7n + 1 int32 words with coordinates, 5n + 1 with the matrix.
template <int BLOCK, int PP, class Dist>
__device__ __forceinline__ void cluster_two_opt_tour(int32_t* __restrict__ tour, int n, const Dist dist, const int2* __restrict__ ct,
                                     int kp, int k, int max_iter, int G, int g, uint64_t* __restrict__ ws,
                                     int64_t* __restrict__ stats, double* __restrict__ cost, int32_t* lds,
                                     uint64_t* red, int64_t* red64, uint64_t* s_best, int P, uint64_t* ws0,
                                     double* __restrict__ cost_all, int32_t* __restrict__ best) {
  constexpr bool XY = __is_same(Dist, IntCoordDist);
  int4* s_nd = reinterpret_cast<int4*>(lds);  // n
  int32_t* s_tour = lds + 4 * n;              // n + 1 (s_tour[n] mirrors s_tour[0])
  int2* s_xy = reinterpret_cast<int2*>(lds + 4 * n + ((n + 2) & ~1));  // n, with coordinates
  const int tid = threadIdx.x;
  const int total = n * k;
  const int base = g * BLOCK + tid;
  const int stride = G * BLOCK;
  // first pair index of this thread's wave: pair slots m with wave_base + m * stride >= total hold
  // no pair in any lane of the wave and are skipped with a wave-uniform branch
  const int wave_base = __builtin_amdgcn_readfirstlane(g * BLOCK + (tid & ~(kWave - 1)));
  uint32_t ac[PP];
  int dac[PP];
  {
    // The table loads, then the tour loads, go out before anything waits: the setup pays one
    // global round trip for both, and one more for the edge costs below.
    int2 e[PP];
    int pa[PP];
    const int qb = stride / k, rb = stride - qb * k;
    int a = base / k, r = base - a * k;
#pragma unroll
    for (int m = 0; m < PP; ++m) {
      pa[m] = a;
      const int p = base + m * stride;
      e[m] = ct[(size_t)(p < total ? a : 0) * kp + (p < total ? r : 0)];
      a += qb;
      r += rb;
      if (r >= k) {
        r -= k;
        ++a;
      }
    }
    load_tour_lds<BLOCK>(tour, n, s_tour);
    if constexpr (XY)
      for (int x = tid; x < n; x += BLOCK) s_xy[x] = dist.xy[x];
#pragma unroll
    for (int m = 0; m < PP; ++m) {
      const int p = base + m * stride;
      const bool ok = p < total && e[m].x >= 0 && e[m].x < n;
      ac[m] = ok ? ((uint32_t)pa[m] | ((uint32_t)e[m].x << 16)) : 0xffffffffu;
      dac[m] = ok ? e[m].y : 0;
    }
  }
  __syncthreads();
  if constexpr (XY) {
    for (int x = tid; x < n; x += BLOCK) {
      const int u = s_tour[x], v = s_tour[x + 1];
      const int2 pu = s_xy[u], pv = s_xy[v];
      s_nd[u] = make_int4(x, euc2d_nint_i32(pu.x, pu.y, pv.x, pv.y), pv.x, pv.y);
    }
  } else {
    build_node_records<BLOCK>(s_nd, s_tour, n, dist);
  }
  __syncthreads();

  const int rows = n - 2;
  int64_t my_deltas = 0;
  int moves = 0, passes = 0;
  bool failed = false;
  int zlo = 1, zhi = 0;  // positions whose edge cost the last move left stale
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    // The stale edge costs' first refresh round (positions zlo + tid + q * BLOCK, q < ZU) is
    // loaded first, then D[succ a][succ c] of every pair, and only then are the refreshed costs
    // stored: the store waits for the refresh loads alone, so both sets of loads are in flight
    // together and a pass waits on global memory once.  The deltas read the costs after the
    // barrier.  Segments longer than ZU * BLOCK finish in refresh_edge_costs.
    constexpr int ZU = 4;
    const bool stale = !XY && zlo <= zhi;
    int zu[ZU], zd[ZU];
    if (stale) {
#pragma unroll
      for (int q = 0; q < ZU; ++q) {
        const int x = zlo + tid + q * BLOCK;
        const bool in = x <= zhi;
        const int u = s_tour[in ? x : 0];
        zd[q] = dist(u, s_tour[in ? x + 1 : 1]);
        zu[q] = in ? u : -1;
      }
    }
    int gd[PP], pi[PP], pj[PP];
#pragma unroll
    for (int m = 0; m < PP; ++m) {
      pi[m] = pj[m] = gd[m] = 0;
      if (wave_base + m * stride >= total) continue;  // no lane of this wave holds pair m
      const uint32_t w = ac[m];
      const int a = (int)(w & 0xffffu), c = (int)(w >> 16);
      if constexpr (XY) {
        // the whole records: the edge costs come with them (gd holds the delta's part from them)
        const int4 na = s_nd[w == 0xffffffffu ? 0 : a];
        const int4 nc = s_nd[w == 0xffffffffu ? 0 : c];
        pi[m] = na.x;
        pj[m] = nc.x;
        gd[m] = euc2d_nint_i32(na.z, na.w, nc.z, nc.w) - na.y - nc.y;
      } else {
        const int2 na = *reinterpret_cast<const int2*>(&s_nd[w == 0xffffffffu ? 0 : a]);
        const int2 nc = *reinterpret_cast<const int2*>(&s_nd[w == 0xffffffffu ? 0 : c]);
        pi[m] = na.x;
        pj[m] = nc.x;
        gd[m] = dist(na.y, nc.y);
      }
    }
    if (stale) {
#pragma unroll
      for (int q = 0; q < ZU; ++q)
        if (zu[q] >= 0) s_nd[zu[q]].z = zd[q];
      if (zhi - zlo >= ZU * BLOCK) refresh_edge_costs<BLOCK>(s_nd, s_tour, zlo + ZU * BLOCK, zhi, dist);
      __syncthreads();
      zlo = 1;
      zhi = 0;
    }
    uint64_t best = kNoMove;
#pragma unroll
    for (int m = 0; m < PP; ++m) {
      if (wave_base + m * stride >= total) continue;
      const uint32_t w = ac[m];
      const int a = (int)(w & 0xffffu), c = (int)(w >> 16);
      const int i = pi[m], j = pj[m];
      const int lim = (i == 0) ? n - 1 : n;
      const bool ok = (w != 0xffffffffu) & ((unsigned)i < (unsigned)rows) & (j > i + 1) & (j < lim);
      const int delta = XY ? dac[m] + gd[m]
                           : dac[m] + gd[m] - s_nd[w == 0xffffffffu ? 0 : a].z - s_nd[w == 0xffffffffu ? 0 : c].z;
      my_deltas += ok;
      if (ok & (delta < 0)) {
        const uint32_t r = (uint32_t)(base + m * stride - a * k);
        const uint64_t key = pack_key(delta, (((uint32_t)i * (uint32_t)k + r) << kRegJBits) | (uint32_t)j);
        best = key < best ? key : best;
      }
    }
    // Only wave 0 needs the workgroup's minimum (it publishes and polls): one barrier, not two.
    // red[] was last read by wave 0 before the previous pass's s_best barrier.
    best = wave_min_u64(best);
    if ((tid & (kWave - 1)) == 0) red[tid / kWave] = best;
    __syncthreads();
    uint64_t* err = ws + 2 * kClusterMaxG * kClusterLine;
    const uint64_t tag = ((it >> 1) & 1) ? 0ull : kClusterTag;
    if (tid < kWave) {
      const int lane = tid;
      best = wave_min_u64(lane < BLOCK / kWave ? red[lane] : kNoMove);
      if (lane == 0) store_relaxed(ws + g * kClusterLine + (it & 1), (best == kNoMove ? kClusterNone : best) | tag);
      bool fail = false;
      uint64_t w = kClusterNone | tag;
      for (unsigned spin = 0;; ++spin) {
        // Every 64th poll (the first included) also reads the error word, in flight with the keys.
        const bool check = (spin & 63) == 0;
        const uint64_t e = check ? load_relaxed(err) : 0ull;
        if (lane < G) w = load_relaxed(ws + lane * kClusterLine + (it & 1));
        if (check && (e != 0 || spin >= kClusterSpinLimit)) {
          fail = true;
          break;
        }
        if (__all((w & kClusterTag) == tag)) break;
        __builtin_amdgcn_s_sleep(1);
      }
      const uint64_t key = wave_min_u64(w & ~kClusterTag);
      if (lane == 0) {
        if (fail) store_relaxed(err, 1ull);
        s_best[0] = fail ? kClusterFail : (key == kClusterNone ? kNoMove : key);
      }
    }
    __syncthreads();
    const uint64_t all = s_best[0];
    if (all == kClusterFail) {
      failed = true;
      break;
    }
    if (all == kNoMove) break;
    const uint32_t idx = key_index(all);
    const int i = (int)((idx >> kRegJBits) / (uint32_t)k);
    const int j = (int)(idx & ((1u << kRegJBits) - 1u));
    const int half = (j - i) >> 1;
    for (int t = tid; t < half; t += BLOCK) {
      const int lo = i + 1 + t, hi = j - t;
      const int32_t tmp = s_tour[lo];
      s_tour[lo] = s_tour[hi];
      s_tour[hi] = tmp;
    }
    __syncthreads();
    for (int x = i + tid; x <= j; x += BLOCK) {
      const int u = s_tour[x], v = s_tour[x + 1];
      if constexpr (XY) {
        const int2 pu = s_xy[u], pv = s_xy[v];
        s_nd[u] = make_int4(x, euc2d_nint_i32(pu.x, pu.y, pv.x, pv.y), pv.x, pv.y);
      } else {
        *reinterpret_cast<int2*>(&s_nd[u]) = make_int2(x, v);  // its cost is refreshed next pass
      }
    }
    __syncthreads();
    if constexpr (!XY) {
      zlo = i;
      zhi = j;
    }
    ++moves;
  }
  if constexpr (!XY) {
    if (!failed && g == 0 && cost != nullptr && zlo <= zhi) {
      refresh_edge_costs<BLOCK>(s_nd, s_tour, zlo, zhi, dist);
      __syncthreads();
    }
  }
  if (!failed && g == 0) {
    for (int x = tid; x < n; x += BLOCK) tour[x] = s_tour[x];
    if (cost != nullptr) {
      int64_t part = 0;
      for (int x = tid; x < n; x += BLOCK) part += XY ? s_nd[x].y : s_nd[x].z;
      const int64_t tot = block_sum_i64(part, red64);
      if (tid == 0) store_relaxed(reinterpret_cast<uint64_t*>(cost), (uint64_t)__double_as_longlong((double)tot));
    }
  }
  if (best != nullptr && cost_all != nullptr && g == 0) {
    uint64_t* count = ws0 + 2 * kClusterMaxG * kClusterLine + 2;
    if (tid == 0) {
      __builtin_amdgcn_s_waitcnt(0);  // this tour's cost is stored before it is counted
      const uint64_t before = __hip_atomic_fetch_add(count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_best[1] = before == (uint64_t)(P - 1) ? 1ull : 0ull;
    }
    __syncthreads();
    if (s_best[1] != 0 && tid < kWave) {
      // The last tour takes the first minimum of the P costs, one wave loading them together.
      uint64_t lo = ~0ull;
      int arg = 0;
      for (int t0 = 0; t0 < P; t0 += kWave) {
        const int t = t0 + tid;
        const uint64_t key =
            t < P ? argmin_bits(__longlong_as_double((long long)load_relaxed(reinterpret_cast<uint64_t*>(cost_all + t))))
                  : ~0ull;
        const uint64_t m = wave_min_u64(key);
        const int first = __builtin_ctzll(__ballot(t < P && key == m));
        if (t0 == 0 || m < lo) {
          lo = m;
          arg = t0 + first;
        }
      }
      if (tid == 0) {
        best[0] = arg;
        store_relaxed(count, 0ull);
      }
    }
  }
  if (stats != nullptr) {
    const int64_t tot = block_sum_i64(my_deltas, red64);
    if (tid == 0) {
      atomicAdd((unsigned long long*)&stats[0], (unsigned long long)tot);
      if (g == 0) {
        atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
        atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
      }
    }
  }
  if (tid == 0) {
    uint64_t* dep = ws + 2 * kClusterMaxG * kClusterLine + 1;
    const uint64_t done = __hip_atomic_fetch_add(dep, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (done == (uint64_t)(G - 1)) {
      for (int w = 0; w < G; ++w) {
        store_relaxed(ws + w * kClusterLine, 0ull);
        store_relaxed(ws + w * kClusterLine + 1, 0ull);
      }
      store_relaxed(dep, 0ull);
    }
  }
}

}  // namespace tspmi
