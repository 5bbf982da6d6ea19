// Population-batched 2-opt for cyclic TSP tours (gfx950).
//
// One workgroup refines one tour for up to `max_iter` passes (or to its local optimum) in a
// single launch.  The tour, its edge costs edge[x] = D[t_x, t_{x+1}] and (candidate mode) the
// inverse map pos[node] live in LDS; distance rows are gathered from the L2-resident matrix.
//
// Semantics follow the reference CPU scan (src/algorithms/strategies/local_search/two_opt.py):
//   full neighbourhood : pairs (i, j), 0 <= i <= n-3, i+2 <= j <= n-1, (0, n-1) excluded,
//                        scan order (i asc, j asc)                        (two_opt.py:158-201)
//   candidate lists    : pairs (i, r), j = pos[N[t_i, r]], kept iff j > i+1 and j < n and
//                        not (i == 0 and j == n-1), scan order (i asc, r asc), duplicates kept
//                                                                          (two_opt.py:347-367)
//   delta              : D[a,c] + D[b,d] - D[a,b] - D[c,d]                 (two_opt.py:237-245)
//   best               : first minimum in scan order, applied iff < 0      (two_opt.py:386-440)
//   first              : first pair in scan order with delta < 0           (two_opt.py:442-489)
//   move               : reverse positions i+1 .. j                        (two_opt.py:438)
// Exact tie-breaking comes from packing (delta, scan index) into one 64-bit key whose unsigned
// minimum is the reference's choice.
#include <mutex>

#include "two_opt_dev.hpp"

namespace tspmi {

// Steepest scan of the full neighbourhood with a diagonal lane mapping.
//
// Lane L of wave w owns, in slot q, the j offset j - i = 2 + off(w, q) + L of every row i, where
// the 64-lane groups off(w, q) cover 0 .. 256*QMAX - 1 (diag_slot_offset).  Moving from row i to
// row i+1 shifts each owned j by one, so the lane's second gather of row i, D[t_{i+1}, t_{j+1}], is
// exactly its first gather of row i+1, D[t_{i+1}, t_{j'}] with j' = j+1: it is carried in a register
// and every delta costs one distance gather instead of two.  The gathers of row i+1 are issued
// before row i is evaluated (one-row software pipeline).  The loop body is branch-free so every
// load is unconditional and the pipeline really overlaps: s_tour and s_edge carry `pad` zero
// entries past their ends (node 0 is a valid gather target; the edge padding is -2^30), so lanes
// past the row end compute a large positive delta that never wins.  The distance-row base is wave-uniform and the gather
// uses a 32-bit unsigned byte offset.  s_tour[n] mirrors s_tour[0], so the successor of position
// n-1 needs no wrap test.
//
// Row i has pairs only up to j = n-1, so slot q of wave w holds a pair only while
// i < (n-2) - off(w, q).  The rows are walked in QMAX phases: phase QL evaluates the QL lowest
// slots of the wave, up to the row where slot QL-1 runs out, so no slot is evaluated past the
// last row in which one of its lanes is admissible (at n = 1000: 530,432 slot evaluations per
// pass for 498,500 pairs, against 923,648 when every slot runs to the wave's last row).  The
// snake order of off() over the waves gives every wave the same number of live slot rows.
//
// The scan keeps only each lane's smallest delta per slot; the row of the workgroup's smallest
// delta is recovered afterwards by walking the diagonals that reached it (scan_full_best_diag),
// which restores the lowest-(i, j) tie-break of the reference's scan order.  Per slot and row the
// bookkeeping is then one three-operand add (delta), one subtraction (the next row's carry, with
// its edge cost folded in) and half a three-operand minimum, and the LDS reads of a block of rows
// share one address per slot (DiagScan::Lanes): 740 VALU instructions per 16-row period of four
// slots against 889 when the row of each slot minimum was tracked and every address recomputed.
//
// With MIX (EUC_2D instances, coordinates staged in LDS as s_xy), 13 rows of every 16 (all but
// rows 0, 6 and 11) compute their new distances from the coordinates instead of gathering them:
// with the FP64 kind by the function that built the matrix (the same values); with the integer
// kind by euc2d_floor_f32, a lower bound at most 1 below the matrix entry, in six instructions
// (the scan then settles the pass's best move exactly, scan_full_best_diag).  Those rows put no
// work on the texture addresser, whose 64-lane random gathers (32 cycles each) bind the
// matrix-only scan.  Measured per n = 1000, P = 1024 launch to the local optimum (585 ms with no
// coordinate rows): FP64 form, 3 rows in 8: 430-440 ms; exact integer form (14 instructions), k
// rows in 8 for k = 3..6: 391 / 320 / 313 / 342 ms, 9 rows in 16: 308 ms; with the bookkeeping
// cut to a minimum per slot and shared addresses: 8 / 9 / 10 / 11 rows in 16: 304 / 277 / 276 /
// 286 ms; with lower-bound rows (integer coordinates in LDS): 9 / 10 / 11 / 12 / 13 rows: 265 /
// 237 / 222 / 224 / 226 ms; with those coordinates held as exact FP32 (no conversions): 10 / 11 /
// 12 / 13 / 14 rows: 232 / 206 / 198 / 197 / 199 ms.
constexpr int kMixPeriod = 16;
constexpr int kMixRows = 0b1111101111011110;

// Coordinate forms of the MIX rows: FP64 coordinates indexed by node with the FP64 distance, or,
// when every coordinate is an integer and each axis spans at most kIntCoordSpan, the coordinates
// minus the box's low corner as exact FP32 pairs kept in tour order (entry x holds the node at
// position x, so a lane's read is contiguous with its neighbours' and needs no lookup through the
// tour).  dist is the exact distance (edge costs, the move's new edges; the integer form for the
// integer kind); scan_dist is what the scan's coordinate rows use, at most kSlack below dist: the
// integer kind takes the truncated FP32 root, so its deltas are lower bounds within 2 of the exact
// ones (two distances per delta).
struct XyF64 {
  using T = double2;
  static constexpr bool kTourOrder = false;
  static constexpr int kSlack = 0;
  static __device__ __forceinline__ int dist(T a, T b) { return euc2d_nint(a.x, a.y, b.x, b.y); }
  static __device__ __forceinline__ int scan_dist(T a, T b) { return dist(a, b); }
};
// Integer coordinates minus the bounding box's low corner, as FP32 (exact: 0 .. kIntCoordSpan).
struct XyInt {
  using T = float2;
  static constexpr bool kTourOrder = true;
  static constexpr int kSlack = 2;
  static __device__ __forceinline__ T stage(double2 p, int lo_x, int lo_y) {
    return make_float2((float)(p.x - lo_x), (float)(p.y - lo_y));
  }
  static __device__ __forceinline__ int dist(T a, T b) {
    return euc2d_nint_i32((int)a.x, (int)a.y, (int)b.x, (int)b.y);
  }
  static __device__ __forceinline__ int scan_dist(T a, T b) { return euc2d_floor_f32(a.x, a.y, b.x, b.y); }
};

// Delta of pairs the scan must not pick (row 0's pair (0, n-1)), and the start of each slot
// minimum: far above every real delta, so such a lane is never a candidate.
constexpr int kFarDelta = 0x3fffffff;

// int32 words of LDS before the tour array in two_opt_mix_kernel: n node-indexed double2, or
// n + 1 + pad tour-ordered float2 (the tour's own length with its mirror entry and padding).
__host__ __device__ constexpr int mix_coord_words(int n, int pad) {
  return ((4 * n > 2 * (n + 1 + pad) ? 4 * n : 2 * (n + 1 + pad)) + 3) & ~3;
}

__device__ __forceinline__ int diag_slot_offset(int w, int q) {
  return kWave * (kTwoOptWaves * q + ((q & 1) ? (kTwoOptWaves - 1 - w) : w));
}

// BYTES: the tour array holds byte offsets 4 * node (two_opt_mix_kernel), so a gather's column
// offset is read ready-made and a row's base is D + offset * ld bytes.
template <int QMAX, bool MIX = false, class XY = XyF64, bool BYTES = false>
struct DiagScan {
  using T = typename XY::T;
  const int32_t* s_tour;
  const int32_t* s_edge;
  const T* s_xy;  // MIX only
  const int32_t* __restrict__ D;
  int ld, n, wid;
  int jo_even, jo_odd;  // j - i of the lane in slots 0 and 1; slot q adds 512 * (q >> 1)
  // carry: D[t_i, t_j] - e_j of the row about to be evaluated; gcur: D[t_{i+1}, t_{j+1}];
  // bd: the slot's smallest delta (its row is found after the scan).
  int carry[QMAX], gcur[QMAX], bd[QMAX];

  static constexpr int kSlotStride = 2 * kTwoOptBlock;
  // node id and distance row of a tour entry
  static __device__ __forceinline__ int node(int e) { return BYTES ? (int)((uint32_t)e >> 2) : e; }
  __device__ __forceinline__ const int32_t* row_of(int e) const {
    return BYTES ? reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(D) + (size_t)(uint32_t)e * ld)
                 : D + (size_t)e * ld;
  }
  static __device__ __forceinline__ int at(const int32_t* __restrict__ row, int e) {
    return BYTES ? *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(row) + (uint32_t)e) : row_at(row, e);
  }
  __device__ __forceinline__ int jo(int q) const { return ((q & 1) ? jo_odd : jo_even) + (q >> 1) * kSlotStride; }

  // The lane's LDS positions at row i, per slot: row i + R reads entry [R] from them, a constant
  // that the LDS instruction takes as its immediate offset, so a block of rows shares one address
  // computation per array and slot.  The slot's index is made opaque first: otherwise the
  // compiler pairs the reads of slots q and q + 2 (512 entries apart) into one two-address LDS
  // instruction, whose offsets are multiples of 64 entries, and then spends a VALU add per pair
  // and row on the address.
  struct Lanes {
    const int32_t* tour[QMAX];  // &s_tour[i + jo + 2]
    const int32_t* edge[QMAX];  // &s_edge[i + jo + 1]
    const T* xy[QMAX];          // &s_xy[i + jo + 2] (tour-ordered coordinates)
  };
  template <int QL>
  __device__ __forceinline__ Lanes lanes(int i) const {
    Lanes b;
#pragma unroll
    for (int q = 0; q < QL; ++q) {
      const int x = opaque_i32(i + jo(q));
      b.tour[q] = s_tour + x + 2;
      b.edge[q] = s_edge + x + 1;
      if constexpr (MIX && XY::kTourOrder) b.xy[q] = s_xy + x + 2;
    }
    return b;
  }

  // Row i + R over the QL lowest slots; FIRST marks row 0, whose pair (0, n-1) is excluded; COORD
  // takes the row's new distances from the coordinates.  delta = carry + gcur - e_i, one
  // three-operand add, and the next row's carry folds in its e_j with one subtraction.
  template <int QL, bool FIRST, bool COORD = false>
  __device__ __forceinline__ void row(const Lanes& b, int i, int R) {
    // MIX: no instruction crosses a row boundary.  Left free, the scheduler hoists the LDS reads
    // and distance arithmetic of later rows of the 16-row period and needs 157 VGPRs (3 waves
    // per SIMD instead of the 4 that four resident workgroups per CU need); with the rows fenced
    // it needs about 90, and the launch runs 281 ms instead of 310 at n = 1000, P = 1024.
    if constexpr (MIX) __builtin_amdgcn_sched_barrier(0);
    const int r = i + R;
    int gnext[QL];
    if constexpr (COORD && XY::kTourOrder) {
      const T pa = s_xy[r + 2];
#pragma unroll
      for (int q = 0; q < QL; ++q) gnext[q] = XY::scan_dist(pa, b.xy[q][R]);
    } else if constexpr (COORD) {
      const T pa = s_xy[node(__builtin_amdgcn_readfirstlane(s_tour[r + 2]))];
#pragma unroll
      for (int q = 0; q < QL; ++q) gnext[q] = XY::scan_dist(pa, s_xy[node(b.tour[q][R])]);
    } else {
      const int32_t* __restrict__ rn = row_of(__builtin_amdgcn_readfirstlane(s_tour[r + 2]));
#pragma unroll
      for (int q = 0; q < QL; ++q) gnext[q] = at(rn, b.tour[q][R]);
    }
    const int neg_eab = opaque_i32(-s_edge[r]);  // one negation per row, then a three-operand add
#pragma unroll
    for (int q = 0; q < QL; ++q) {
      int delta = carry[q] + gcur[q] + neg_eab;
      if (FIRST) delta = (jo(q) < n - 1) ? delta : kFarDelta;
      bd[q] = min(bd[q], delta);
      carry[q] = opaque_i32(gcur[q] - b.edge[q][R]);  // kept: else the sum is reassociated
      gcur[q] = gnext[q];
    }
  }

  // Rows i + R .. i + kMixPeriod - 1 of a MIX period.
  template <int QL, int R = 0>
  __device__ __forceinline__ void mix_rows(const Lanes& b, int i) {
    if constexpr (R < kMixPeriod) {
      row<QL, false, ((kMixRows >> R) & 1) != 0>(b, i, R);
      mix_rows<QL, R + 1>(b, i);
    }
  }

  // Phases QL, QL-1, ..., 1 from row i on.
  template <int QL>
  __device__ __forceinline__ void phases(int i) {
    if constexpr (QL > 0) {
      const int end = (n - 2) - diag_slot_offset(wid, QL - 1);
      if constexpr (MIX) {
        for (; i + kMixPeriod - 1 < end; i += kMixPeriod) mix_rows<QL>(lanes<QL>(i), i);
      }
      for (; i + 1 < end; i += 2) {  // two rows per trip: the carry moves become register renames
        const Lanes b = lanes<QL>(i);
        row<QL, false>(b, i, 0);
        row<QL, false>(b, i, 1);
      }
      if (i < end) {
        row<QL, false>(lanes<QL>(i), i, 0);
        ++i;
      }
      phases<QL - 1>(i);
    }
  }
};

// The smallest exact (delta, scan index) key over lane L's diagonal of slot q of this wave, the
// same in every lane: the wave walks the diagonal 64 rows at a time over exactly the rows the scan
// gave the slot (row 0, then rows 1 .. (n - 2) - off(w, q) - 1) and takes every delta from the
// matrix.  Four blocks of 64 rows per trip, branch-free (indices clamped into the tour, the key
// of an inadmissible lane discarded), so the loads of the four are in flight together.
template <bool BYTES>
__device__ __forceinline__ uint64_t diag_best_key(const int32_t* s_tour, const int32_t* s_edge, int n,
                                                  const int32_t* __restrict__ D, int ld, int wid, int lane, int q,
                                                  int L) {
  const int off = diag_slot_offset(wid, q);
  const int jo = 2 + off + L;
  const int end = (n - 2) - off;
  const int last = end > 1 ? end : 1;
  uint64_t best = kNoMove;
  for (int r0 = 0; r0 < last; r0 += 4 * kWave) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = r0 + u * kWave + lane;
      const int j = i + jo;
      const bool ok = i < last && j <= n - 1 && !(i == 0 && j == n - 1);
      const int ic = ok ? i : 0, jc = ok ? j : 2;  // (0, 2) is a pair of every tour with n >= 4
      const int a = s_tour[ic], b = s_tour[ic + 1], c = s_tour[jc], d = s_tour[jc + 1];
      const int dac = DiagScan<1, false, XyF64, BYTES>::at(
          reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(D) + (BYTES ? (size_t)(uint32_t)a * ld
                                                                                    : (size_t)a * ld * 4)), c);
      const int dbd = DiagScan<1, false, XyF64, BYTES>::at(
          reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(D) + (BYTES ? (size_t)(uint32_t)b * ld
                                                                                    : (size_t)b * ld * 4)), d);
      const int delta = dac + dbd - s_edge[ic] - s_edge[jc];
      const uint64_t key = ok ? pack_key(delta, (uint32_t)i * (uint32_t)n + (uint32_t)j) : kNoMove;
      best = key < best ? key : best;
    }
  }
  return wave_min_u64(best);
}

// The pass's best move as a packed (delta, scan index) key, the same in every thread of the
// workgroup (kNoMove if no delta is negative).  The scan keeps only each slot's smallest delta
// (one instruction per pair instead of a compare and two selects that would also track its row),
// and with the integer coordinate kind its coordinate rows use lower-bound distances (kSlack): a
// scanned delta d~ satisfies d - kSlack <= d~ <= d.  With m the workgroup's smallest d~: every
// exact delta is >= m, so m >= 0 means no improving move; otherwise a pair holding the exact
// minimum d* <= (the exact delta of m's pair) <= m + kSlack has d~ <= d* <= m + kSlack, so every
// lane whose slot minimum is within m + kSlack has its diagonal walked again with exact deltas by
// its wave (diag_best_key) -- normally one lane in the workgroup -- and the smallest key among
// them is the reference's first minimum in scan order.  red holds 2 * waves entries.
template <int QMAX, bool MIX = false, class XY = XyF64, bool BYTES = false>
__device__ __forceinline__ uint64_t scan_full_best_diag(const int32_t* s_tour, const int32_t* s_edge, int n,
                                                        const int32_t* __restrict__ D, int ld, int wid, int lane,
                                                        uint64_t* red, const typename XY::T* s_xy = nullptr) {
  DiagScan<QMAX, MIX, XY, BYTES> sc;
  sc.s_tour = s_tour;
  sc.s_edge = s_edge;
  sc.s_xy = s_xy;
  sc.D = D;
  sc.ld = ld;
  sc.n = n;
  sc.wid = wid;
  sc.jo_even = 2 + diag_slot_offset(wid, 0) + lane;
  sc.jo_odd = 2 + diag_slot_offset(wid, 1) + lane;
  {
    const int32_t* __restrict__ r0 = sc.row_of(__builtin_amdgcn_readfirstlane(s_tour[0]));
    const int32_t* __restrict__ r1 = sc.row_of(__builtin_amdgcn_readfirstlane(s_tour[1]));
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int j = sc.jo(q);
      sc.carry[q] = sc.at(r0, s_tour[j]) - s_edge[j];
      sc.gcur[q] = sc.at(r1, s_tour[j + 1]);
      sc.bd[q] = kFarDelta;
    }
  }
  sc.template row<QMAX, true>(sc.template lanes<QMAX>(0), 0, 0);
  sc.template phases<QMAX>(1);
  int m = sc.bd[0];
#pragma unroll
  for (int q = 1; q < QMAX; ++q) m = min(m, sc.bd[q]);
  m = wave_min_i32(m);
  // red[0 .. waves) is last read in the previous pass, before barriers of the move
  int* red_m = reinterpret_cast<int*>(red);
  if (lane == 0) red_m[wid] = m;
  __syncthreads();
  int dmin = red_m[0];
  for (int w = 1; w < kTwoOptWaves; ++w) dmin = min(dmin, red_m[w]);
  if (dmin >= 0) return kNoMove;  // the same in every thread
  const int limit = dmin + (MIX ? XY::kSlack : 0);
  uint32_t cand_slots = 0;
#pragma unroll
  for (int q = 0; q < QMAX; ++q) cand_slots |= (sc.bd[q] <= limit ? 1u : 0u) << q;
  uint64_t best = kNoMove;
  for (int q = 0; q < QMAX; ++q) {
    uint64_t mask = __ballot((cand_slots >> q) & 1u);
    while (mask != 0) {
      const int L = __ffsll((unsigned long long)mask) - 1;
      mask &= mask - 1;
      const uint64_t key = diag_best_key<BYTES>(s_tour, s_edge, n, D, ld, wid, lane, q, L);
      best = key < best ? key : best;
    }
  }
  if (lane == 0) red[kTwoOptWaves + wid] = best;
  __syncthreads();
  for (int w = 0; w < kTwoOptWaves; ++w) best = red[kTwoOptWaves + w] < best ? red[kTwoOptWaves + w] : best;
  // The exact minimum may be >= 0 although a lower bound was negative: then there is no move.
  return best < pack_key(0, 0) ? best : kNoMove;
}

// Large-n variant of the diagonal scan: 16 waves per tour and clamped LDS indices instead of
// padding (the padding would not fit next to a 10^4-city tour), with the explicit bound test.
constexpr int kWideBlock = 1024;
constexpr int kWideWaves = kWideBlock / kWave;

template <int QMAX>
__device__ __forceinline__ uint64_t scan_full_best_wide(const int32_t* s_tour, const int32_t* s_edge, int n,
                                                        const int32_t* __restrict__ D, int ld, int wid, int lane) {
  const int rows = n - 2;
  const int jq0 = 2 + wid * kWave + lane;
  int carry[QMAX], gcur[QMAX], gnext[QMAX], bd[QMAX], bi[QMAX];
  {
    const int32_t* __restrict__ r0 = D + (size_t)__builtin_amdgcn_readfirstlane(s_tour[0]) * ld;
    const int32_t* __restrict__ r1 = D + (size_t)__builtin_amdgcn_readfirstlane(s_tour[1]) * ld;
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int j = jq0 + q * kWideBlock;
      carry[q] = row_at(r0, s_tour[min(j, n)]);
      gcur[q] = row_at(r1, s_tour[min(j + 1, n)]);
      bd[q] = 0;
      bi[q] = 0;
    }
  }
  for (int i = 0; i < rows; ++i) {
    if (i + 2 + kWave * wid >= n) break;
    const int32_t* __restrict__ rn = D + (size_t)__builtin_amdgcn_readfirstlane(s_tour[i + 2]) * ld;
#pragma unroll
    for (int q = 0; q < QMAX; ++q) gnext[q] = row_at(rn, s_tour[min(i + jq0 + 2 + q * kWideBlock, n)]);
    const int eab = s_edge[i];
    const int lim = (i == 0) ? n - 1 : n;
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int j = i + jq0 + q * kWideBlock;
      const int delta = carry[q] + gcur[q] - eab - s_edge[min(j, n - 1)];
      const bool take = (j < lim) & (delta < bd[q]);
      bd[q] = take ? delta : bd[q];
      bi[q] = take ? i : bi[q];
      carry[q] = gcur[q];
      gcur[q] = gnext[q];
    }
  }
  uint64_t best = kNoMove;
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    if (bd[q] < 0) {
      const uint32_t r = (uint32_t)bi[q];
      const uint64_t key = pack_key(bd[q], r * (uint32_t)n + r + (uint32_t)(jq0 + q * kWideBlock));
      best = key < best ? key : best;
    }
  }
  return best;
}

template <int QMAX>
__global__ __launch_bounds__(kWideBlock) void two_opt_wide_kernel(int32_t* __restrict__ tours, int n,
                                                                  const int32_t* __restrict__ D, int ld, int max_iter,
                                                                  int64_t* __restrict__ stats) {
  extern __shared__ int32_t lds[];
  __shared__ uint64_t red[kWideWaves];
  int32_t* s_tour = lds;          // n + 1 (s_tour[n] mirrors s_tour[0])
  int32_t* s_edge = lds + n + 1;  // n
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = wave_id();
  int32_t* tour = tours + (size_t)blockIdx.x * n;
  for (int x = tid; x < n; x += kWideBlock) {
    const int v = clamp_id(tour[x], ld);
    s_tour[x] = v;
    if (x == 0) s_tour[n] = v;
  }
  __syncthreads();
  for (int x = tid; x < n; x += kWideBlock) s_edge[x] = D[(size_t)s_tour[x] * ld + s_tour[x + 1]];
  __syncthreads();
  int moves = 0, passes = 0;
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    uint64_t best = scan_full_best_wide<QMAX>(s_tour, s_edge, n, D, ld, wid, lane);
    best = block_min_u64(best, red);
    if (best == kNoMove) break;
    const uint32_t idx = key_index(best);
    const int i = (int)(idx / (uint32_t)n);
    const int j = (int)(idx - (uint32_t)i * (uint32_t)n);
    __syncthreads();
    apply_move(s_tour, s_edge, nullptr, i, j, n, D, ld);
    ++moves;
  }
  for (int x = tid; x < n; x += kWideBlock) tour[x] = s_tour[x];
  if (stats != nullptr && tid == 0) {
    atomicAdd((unsigned long long*)&stats[0], (unsigned long long)((int64_t)passes * n * (n - 3) / 2));
    atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
    atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
  }
}

// Row-staged variant for rows too long for the L1 (n > ~2000): the diagonal scan reads the new
// row D[t_{i+2}][*] at n random columns per step; here the whole row is first copied into LDS with
// coalesced loads (issued two steps ahead into registers), so the random accesses hit LDS and
// every 128-byte line of the row is fetched once per step.  Two workgroup barriers per step.
// Needs (3n + 1) int32 of LDS and n - 2 <= QMAX * 1024 + 1 (QMAX <= 10 keeps it within 128 VGPRs).
template <int E>
__device__ __forceinline__ void fetch_row(int (&r)[E], const int32_t* __restrict__ D, int ld, int node, int n) {
  const int32_t* __restrict__ row = D + (size_t)node * ld;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = threadIdx.x + e * kWideBlock;
    r[e] = x < n ? row[x] : 0;
  }
}

template <int E>
__device__ __forceinline__ void stage_row(const int (&r)[E], int32_t* s_row, int n) {
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = threadIdx.x + e * kWideBlock;
    if (x < n) s_row[x] = r[e];
  }
}

template <int QMAX>
__device__ __forceinline__ uint64_t scan_full_best_staged(const int32_t* s_tour, const int32_t* s_edge, int32_t* s_row,
                                                          int n, const int32_t* __restrict__ D, int ld, int wid,
                                                          int lane) {
  constexpr int E = QMAX + 1;
  const int rows = n - 2;
  const int jq0 = 2 + wid * kWave + lane;
  int carry[QMAX], gcur[QMAX], bd[QMAX], bi[QMAX];
  {
    const int32_t* __restrict__ r0 = D + (size_t)__builtin_amdgcn_readfirstlane(s_tour[0]) * ld;
    const int32_t* __restrict__ r1 = D + (size_t)__builtin_amdgcn_readfirstlane(s_tour[1]) * ld;
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int j = jq0 + q * kWideBlock;
      carry[q] = row_at(r0, s_tour[min(j, n)]);
      gcur[q] = row_at(r1, s_tour[min(j + 1, n)]);
      bd[q] = 0;
      bi[q] = 0;
    }
  }
  int ra[E], rb[E];
  fetch_row(ra, D, ld, __builtin_amdgcn_readfirstlane(s_tour[2]), n);
  fetch_row(rb, D, ld, __builtin_amdgcn_readfirstlane(s_tour[min(3, n)]), n);
  auto step = [&](int i) {
    if (i + 2 + kWave * wid >= n) return;  // every column of this wave lies past the tour
    const int eab = s_edge[i];
    const int lim = (i == 0) ? n - 1 : n;
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
      const int j = i + jq0 + q * kWideBlock;
      const int gnext = s_row[s_tour[min(j + 2, n)]];
      const int delta = carry[q] + gcur[q] - eab - s_edge[min(j, n - 1)];
      const bool take = (j < lim) & (delta < bd[q]);
      bd[q] = take ? delta : bd[q];
      bi[q] = take ? i : bi[q];
      carry[q] = gcur[q];
      gcur[q] = gnext;
    }
  };
  for (int i = 0; i < rows; i += 2) {
    __syncthreads();
    stage_row(ra, s_row, n);
    if (i + 4 < n) fetch_row(ra, D, ld, __builtin_amdgcn_readfirstlane(s_tour[i + 4]), n);
    __syncthreads();
    step(i);
    if (i + 1 >= rows) break;
    __syncthreads();
    stage_row(rb, s_row, n);
    if (i + 5 < n) fetch_row(rb, D, ld, __builtin_amdgcn_readfirstlane(s_tour[i + 5]), n);
    __syncthreads();
    step(i + 1);
  }
  uint64_t best = kNoMove;
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    if (bd[q] < 0) {
      const uint32_t r = (uint32_t)bi[q];
      const uint64_t key = pack_key(bd[q], r * (uint32_t)n + r + (uint32_t)(jq0 + q * kWideBlock));
      best = key < best ? key : best;
    }
  }
  return best;
}

template <int QMAX>
__global__ __launch_bounds__(kWideBlock) void two_opt_staged_kernel(int32_t* __restrict__ tours, int n,
                                                                    const int32_t* __restrict__ D, int ld,
                                                                    int max_iter, int64_t* __restrict__ stats) {
  extern __shared__ int32_t lds[];
  __shared__ uint64_t red[kWideWaves];
  int32_t* s_tour = lds;          // n + 1 (s_tour[n] mirrors s_tour[0])
  int32_t* s_edge = lds + n + 1;  // n
  int32_t* s_row = s_edge + n;    // n
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = wave_id();
  int32_t* tour = tours + (size_t)blockIdx.x * n;
  for (int x = tid; x < n; x += kWideBlock) {
    const int v = clamp_id(tour[x], n);
    s_tour[x] = v;
    if (x == 0) s_tour[n] = v;
  }
  __syncthreads();
  for (int x = tid; x < n; x += kWideBlock) s_edge[x] = D[(size_t)s_tour[x] * ld + s_tour[x + 1]];
  __syncthreads();
  int moves = 0, passes = 0;
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    uint64_t best = scan_full_best_staged<QMAX>(s_tour, s_edge, s_row, n, D, ld, wid, lane);
    best = block_min_u64(best, red);
    if (best == kNoMove) break;
    const uint32_t idx = key_index(best);
    const int i = (int)(idx / (uint32_t)n);
    const int j = (int)(idx - (uint32_t)i * (uint32_t)n);
    __syncthreads();
    apply_move(s_tour, s_edge, nullptr, i, j, n, D, ld);
    ++moves;
  }
  for (int x = tid; x < n; x += kWideBlock) tour[x] = s_tour[x];
  if (stats != nullptr && tid == 0) {
    atomicAdd((unsigned long long*)&stats[0], (unsigned long long)((int64_t)passes * n * (n - 3) / 2));
    atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
    atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
  }
}

template <int QMAX>
__global__ __launch_bounds__(kTwoOptBlock) void two_opt_kernel(int32_t* __restrict__ tours, int n,
                                                               const int32_t* __restrict__ D, int ld,
                                                               const int32_t* __restrict__ cand, int k, int mode,
                                                               int max_iter, int pad, int64_t* __restrict__ stats) {
  extern __shared__ int32_t lds[];
  __shared__ TwoOptShared sh;
  // s_tour: n + 1 entries (s_tour[n] mirrors s_tour[0]; position 0 never moves) then `pad` zeros;
  // s_edge: n entries then `pad` zeros; s_pos: n entries (candidate mode).
  int32_t* s_tour = lds;
  int32_t* s_edge = lds + n + 1 + pad;
  int32_t* s_pos = (cand != nullptr) ? s_edge + n + pad : nullptr;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = wave_id();
  int32_t* tour = tours + (size_t)blockIdx.x * n;

  for (int x = tid; x < n; x += blockDim.x) {
    int v = tour[x];
    v = v < 0 ? 0 : (v >= ld ? ld - 1 : v);  // ids outside [0, ld) are an invalid input, never a fault
    s_tour[x] = v;
    if (x == 0) s_tour[n] = v;
  }
  // Padding: node 0 past the tour (a valid gather target) and a sentinel edge cost past the edges
  // that keeps every out-of-row delta positive (0 <= D < 2^29, so carry + g - eab + 2^30 > 0).
  for (int x = tid; x < pad; x += blockDim.x) {
    s_tour[n + 1 + x] = 0;
    s_edge[n + x] = -(1 << 30);
  }
  __syncthreads();
  for (int x = tid; x < n; x += blockDim.x) {
    const int a = s_tour[x];
    s_edge[x] = D[(size_t)a * ld + s_tour[wrap_next(x, n)]];
    if (s_pos != nullptr) s_pos[a] = x;
  }
  __syncthreads();

  int64_t my_deltas = 0;
  int moves = 0, passes = 0;
  const int rows = n - 2;  // i in [0, n-3]

  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    uint64_t best = kNoMove;
    if (cand == nullptr && mode == TSPMI_MODE_BEST && QMAX > 0) {
      best = scan_full_best_diag<(QMAX > 0 ? QMAX : 1)>(s_tour, s_edge, n, D, ld, wid, lane, sh.red);
      if (tid == 0) my_deltas += (int64_t)n * (n - 3) / 2;
    } else if (cand == nullptr && mode == TSPMI_MODE_BEST) {
      for (int i = wid; i < rows; i += kTwoOptWaves) {
        const int a = s_tour[i], b = s_tour[i + 1];
        const int eab = s_edge[i];
        const int32_t* __restrict__ ra = D + (size_t)a * ld;
        const int32_t* __restrict__ rb = D + (size_t)b * ld;
        const int lim = (i == 0) ? n - 1 : n;
        const uint32_t base = (uint32_t)i * (uint32_t)n;
        for (int j = i + 2 + lane; j < lim; j += kWave) {
          const int c = s_tour[j];
          const int d = s_tour[wrap_next(j, n)];
          const int delta = ra[c] + rb[d] - eab - s_edge[j];
          if (delta < 0) {
            const uint64_t key = pack_key(delta, base + (uint32_t)j);
            best = key < best ? key : best;
          }
        }
      }
      if (tid == 0) my_deltas += (int64_t)n * (n - 3) / 2;
      best = block_min_u64(best, sh.red);
    } else if (cand == nullptr) {  // first improvement, full neighbourhood
      for (int base_i = 0; base_i < rows; base_i += kTwoOptWaves) {
        const int i = base_i + wid;
        if (i < rows) {
          const int a = s_tour[i], b = s_tour[i + 1];
          const int eab = s_edge[i];
          const int32_t* __restrict__ ra = D + (size_t)a * ld;
          const int32_t* __restrict__ rb = D + (size_t)b * ld;
          const int lim = (i == 0) ? n - 1 : n;
          for (int j0 = i + 2; j0 < lim; j0 += kWave) {
            const int j = j0 + lane;
            bool imp = false;
            if (j < lim) {
              const int c = s_tour[j];
              const int d = s_tour[wrap_next(j, n)];
              imp = (ra[c] + rb[d] - eab - s_edge[j]) < 0;
              ++my_deltas;
            }
            if (imp) best = best < (uint64_t)((uint32_t)i * (uint32_t)n + (uint32_t)j) ? best
                                                                         : (uint64_t)((uint32_t)i * (uint32_t)n + (uint32_t)j);
            if (__any(imp)) break;  // later j in this row cannot be first
          }
        }
        best = block_min_u64(best, sh.red);
        if (best != kNoMove) break;
      }
      if (best != kNoMove) best = pack_key(-1, (uint32_t)best);
    } else {  // candidate lists
      const int total = rows * k;
      if (mode == TSPMI_MODE_BEST) {
        for (int f = tid; f < total; f += blockDim.x) {
          const int i = f / k;
          const int a = s_tour[i];
          const int nb = cand[(size_t)a * k + (f - i * k)];
          if (nb < 0 || nb >= n) continue;
          const int j = s_pos[nb];
          const int lim = (i == 0) ? n - 1 : n;
          if (j <= i + 1 || j >= lim) continue;
          const int b = s_tour[i + 1];
          const int d = s_tour[wrap_next(j, n)];
          const int delta = D[(size_t)a * ld + nb] + D[(size_t)b * ld + d] - s_edge[i] - s_edge[j];
          ++my_deltas;
          if (delta < 0) {
            const uint64_t key = pack_key(delta, (uint32_t)f);
            best = key < best ? key : best;
          }
        }
        best = block_min_u64(best, sh.red);
      } else {
        for (int f0 = 0; f0 < total; f0 += blockDim.x) {
          const int f = f0 + tid;
          if (f < total) {
            const int i = f / k;
            const int a = s_tour[i];
            const int nb = cand[(size_t)a * k + (f - i * k)];
            if (nb >= 0 && nb < n) {
              const int j = s_pos[nb];
              const int lim = (i == 0) ? n - 1 : n;
              if (j > i + 1 && j < lim) {
                const int b = s_tour[i + 1];
                const int d = s_tour[wrap_next(j, n)];
                const int delta = D[(size_t)a * ld + nb] + D[(size_t)b * ld + d] - s_edge[i] - s_edge[j];
                ++my_deltas;
                if (delta < 0) best = (uint64_t)f;
              }
            }
          }
          best = block_min_u64(best, sh.red);
          if (best != kNoMove) break;
        }
        if (best != kNoMove) best = pack_key(-1, (uint32_t)best);
      }
    }

    if (best == kNoMove) break;  // uniform: every thread holds the block minimum
    const uint32_t idx = key_index(best);
    int i, j;
    if (cand == nullptr) {
      i = (int)(idx / (uint32_t)n);
      j = (int)(idx - (uint32_t)i * (uint32_t)n);
    } else {
      i = (int)(idx / (uint32_t)k);
      j = s_pos[cand[(size_t)s_tour[i] * k + (int)(idx - (uint32_t)i * (uint32_t)k)]];
    }
    __syncthreads();  // all threads decoded (i, j) before the tour changes
    apply_move(s_tour, s_edge, s_pos, i, j, n, D, ld);
    ++moves;
  }

  for (int x = tid; x < n; x += blockDim.x) tour[x] = s_tour[x];
  if (stats != nullptr) {
    const int64_t total = block_sum_i64(my_deltas, sh.red64);
    if (tid == 0) {
      atomicAdd((unsigned long long*)&stats[0], (unsigned long long)total);
      atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
      atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
    }
  }
}

// Whether every coordinate is an integer (|v| <= 2^30) and each axis spans at most kIntCoordSpan,
// so that the MIX rows may use the integer distance form; the same answer in every thread.
__device__ __forceinline__ bool int_coords_fit(const double2* __restrict__ xy, int n, int* s_box) {
  const int tid = threadIdx.x;
  if (tid < 4) s_box[tid] = 0x7fffffff;
  __syncthreads();
  int ok = 1, lo_x = 0x7fffffff, lo_y = 0x7fffffff, hi_x = 0x7fffffff, hi_y = 0x7fffffff;  // hi as -max
  for (int x = tid; x < n; x += blockDim.x) {
    const double2 p = xy[x];
    const bool fits = p.x == __builtin_rint(p.x) && p.y == __builtin_rint(p.y) && __builtin_fabs(p.x) <= 1073741824.0 &&
                      __builtin_fabs(p.y) <= 1073741824.0;
    ok &= fits ? 1 : 0;
    if (fits) {
      const int ix = (int)p.x, iy = (int)p.y;
      lo_x = min(lo_x, ix);
      lo_y = min(lo_y, iy);
      hi_x = min(hi_x, -ix);
      hi_y = min(hi_y, -iy);
    }
  }
  atomicMin(&s_box[0], lo_x);
  atomicMin(&s_box[1], hi_x);
  atomicMin(&s_box[2], lo_y);
  atomicMin(&s_box[3], hi_y);
  ok = __syncthreads_and(ok);
  // -s_box[1] - s_box[0] is the x span (both bounds exist once n >= 1)
  return ok && (long)-s_box[1] - s_box[0] <= kIntCoordSpan && (long)-s_box[3] - s_box[2] <= kIntCoordSpan;
}

// Reverse positions i+1 .. j of the tour and of its tour-ordered coordinates, then refresh the
// touched edge costs from the coordinates (s_xy[n] mirrors s_xy[0], so no wrap test).
template <class XY>
__device__ __forceinline__ void apply_move_tour_xy(int32_t* s_tour, typename XY::T* s_xy, int32_t* s_edge, int i,
                                                   int j) {
  const int half = (j - i) >> 1;
  for (int t = threadIdx.x; t < half; t += blockDim.x) {
    const int lo = i + 1 + t, hi = j - t;
    const int32_t tmp = s_tour[lo];
    s_tour[lo] = s_tour[hi];
    s_tour[hi] = tmp;
    const typename XY::T c = s_xy[lo];
    s_xy[lo] = s_xy[hi];
    s_xy[hi] = c;
  }
  __syncthreads();
  for (int x = i + threadIdx.x; x <= j; x += blockDim.x) s_edge[x] = XY::dist(s_xy[x], s_xy[x + 1]);
  __syncthreads();
}

// The matrix read with byte-offset tour entries (4 * node): D[u][v] at byte u * ld + v.
struct ByteMatrixDist {
  const int32_t* __restrict__ D;
  int ld;
  __device__ __forceinline__ int operator()(int u4, int v4) const {
    return *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(D) + (size_t)(uint32_t)u4 * ld +
                                             (uint32_t)v4);
  }
};

// The pass loop of two_opt_mix_kernel with the coordinates staged in LDS in the form XY (s_tour,
// as byte offsets 4 * node, and its padding already in place).
template <int QMAX, class XY>
__device__ __forceinline__ void mix_passes(const double2* __restrict__ xy, typename XY::T* s_xy, int32_t* s_tour,
                                           int32_t* s_edge, int n, int pad, const int32_t* __restrict__ D, int ld,
                                           int max_iter, uint64_t* red, int& moves, int& passes, int lo_x = 0,
                                           int lo_y = 0) {
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = wave_id();
  if constexpr (XY::kTourOrder) {
    for (int x = tid; x < n + 1 + pad; x += kTwoOptBlock) {
      // the mirror entry and the padding (node 0) included
      s_xy[x] = XY::stage(xy[(uint32_t)s_tour[x] >> 2], lo_x, lo_y);
    }
  } else {
    for (int x = tid; x < n; x += kTwoOptBlock) s_xy[x] = xy[x];
  }
  __syncthreads();
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    const uint64_t best =
        scan_full_best_diag<QMAX, true, XY, true>(s_tour, s_edge, n, D, ld, wid, lane, red, s_xy);
    if (best == kNoMove) break;
    const uint32_t idx = key_index(best);
    const int i = (int)(idx / (uint32_t)n);
    const int j = (int)(idx - (uint32_t)i * (uint32_t)n);
    __syncthreads();
    if constexpr (XY::kTourOrder)
      apply_move_tour_xy<XY>(s_tour, s_xy, s_edge, i, j);
    else
      apply_move(s_tour, s_edge, nullptr, i, j, n, ByteMatrixDist{D, ld});
    ++moves;
  }
}

// Full neighbourhood, best improvement, for EUC_2D instances: the diagonal scan with its MIX rows
// from the coordinates, staged in LDS before the tour (mix_coord_words, then the tour and edge
// arrays with their padding).  Same pairs, deltas, tie-break, moves and counters as two_opt_kernel.
// Integer coordinates within kIntCoordSpan per axis are staged as exact FP32 offsets from the box
// corner in tour order and scanned with lower-bound distances; others stay FP64, indexed by node.
template <int QMAX>
__global__ __launch_bounds__(kTwoOptBlock) void two_opt_mix_kernel(int32_t* __restrict__ tours, int n,
                                                                   const int32_t* __restrict__ D, int ld,
                                                                   const double2* __restrict__ xy, int max_iter,
                                                                   int pad, int64_t* __restrict__ stats) {
  extern __shared__ __align__(16) int32_t lds[];
  __shared__ uint64_t red[2 * kTwoOptWaves];
  __shared__ int s_box[4];
  // n + 1 + pad byte offsets 4 * node (the gathers' column offsets as stored), after the coordinates
  int32_t* s_tour = lds + mix_coord_words(n, pad);
  int32_t* s_edge = s_tour + n + 1 + pad;  // n + pad
  const int tid = threadIdx.x;
  int32_t* tour = tours + (size_t)blockIdx.x * n;
  const bool int_xy = int_coords_fit(xy, n, s_box);
  for (int x = tid; x < n; x += kTwoOptBlock) {
    const int v = 4 * clamp_id(tour[x], n);
    s_tour[x] = v;
    if (x == 0) s_tour[n] = v;
  }
  for (int x = tid; x < pad; x += kTwoOptBlock) {
    s_tour[n + 1 + x] = 0;
    s_edge[n + x] = -(1 << 30);
  }
  __syncthreads();
  for (int x = tid; x < n; x += kTwoOptBlock) s_edge[x] = ByteMatrixDist{D, ld}(s_tour[x], s_tour[x + 1]);
  int moves = 0, passes = 0;
  if (int_xy)
    mix_passes<QMAX, XyInt>(xy, reinterpret_cast<float2*>(lds), s_tour, s_edge, n, pad, D, ld, max_iter, red, moves,
                            passes, s_box[0], s_box[2]);
  else
    mix_passes<QMAX, XyF64>(xy, reinterpret_cast<double2*>(lds), s_tour, s_edge, n, pad, D, ld, max_iter, red, moves,
                            passes);
  for (int x = tid; x < n; x += kTwoOptBlock) tour[x] = (int)((uint32_t)s_tour[x] >> 2);
  if (stats != nullptr && tid == 0) {
    atomicAdd((unsigned long long*)&stats[0], (unsigned long long)((int64_t)passes * n * (n - 3) / 2));
    atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
    atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
  }
}

template <int BLOCK, int KP, int UNROLL, int KL>
__global__ __launch_bounds__(BLOCK) void two_opt_cand_kernel(int32_t* __restrict__ tours, int n,
                                                             const int32_t* __restrict__ D, int ld,
                                                             const int2* __restrict__ ct, int k, int max_iter,
                                                             int64_t* __restrict__ stats, double* __restrict__ cost) {
  extern __shared__ int32_t lds[];
  __shared__ uint64_t red[BLOCK / kWave];
  __shared__ int64_t red64[BLOCK / kWave];
  cand_two_opt_tour<BLOCK, KP, UNROLL, KL>(tours + (size_t)blockIdx.x * n, n, MatrixDist{D, ld}, ct, k, max_iter,
                                           stats, lds, red, red64, cost != nullptr ? cost + blockIdx.x : nullptr);
}

// One tour per workgroup of 1024 threads, pairs in registers (reg_two_opt_tour).
template <int PP>
__global__ __launch_bounds__(1024) void two_opt_reg_kernel(int32_t* __restrict__ tours, int n,
                                                           const int32_t* __restrict__ D, int ld,
                                                           const int2* __restrict__ ct, int kp, int k, int max_iter,
                                                           int64_t* __restrict__ stats, double* __restrict__ cost) {
  extern __shared__ __align__(16) int32_t lds[];
  __shared__ uint64_t red[1024 / kWave];
  __shared__ int64_t red64[1024 / kWave];
  reg_two_opt_tour<1024, PP>(tours + (size_t)blockIdx.x * n, n, MatrixDist{D, ld}, ct, kp, k, max_iter, stats,
                             cost != nullptr ? cost + blockIdx.x : nullptr, lds, red, red64);
}

// G workgroups of kClusterBlock threads per tour (cluster_two_opt_tour); tour t is blocks t*G ..
// t*G + G - 1.
constexpr int kClusterBlock = 256;

// XY: distances from int32 coordinates ixy (cluster_two_opt_tour with IntCoordDist), else from D.
template <int PP, bool XY = false>
__global__ __launch_bounds__(kClusterBlock) void two_opt_cluster_kernel(int32_t* __restrict__ tours, int n,
                                                                        const int32_t* __restrict__ D, int ld,
                                                                        const int2* __restrict__ ixy,
                                                                        const int2* __restrict__ ct, int kp, int k,
                                                                        int max_iter, int G, uint64_t* __restrict__ sync,
                                                                        int64_t* __restrict__ stats,
                                                                        double* __restrict__ cost, int P,
                                                                        int32_t* __restrict__ best, int xcds) {
  extern __shared__ __align__(16) int32_t lds[];
  __shared__ uint64_t red[kClusterBlock / kWave];
  __shared__ int64_t red64[kClusterBlock / kWave];
  __shared__ uint64_t s_best[2];  // the pass's move; the last-tour flag
  // Workgroups are handed to the XCDs in turn (workgroup b to XCD b % xcds); with xcds > 1 the G
  // workgroups of a tour are taken from one XCD's share, so a cluster's barrier traffic stays on
  // the L2 of one XCD.  Either mapping is a bijection of blocks onto (tour, member).
  int t, g;
  if (xcds > 1) {
    const int x = blockIdx.x % xcds, slot = blockIdx.x / xcds;
    t = x * (P / xcds) + slot / G;
    g = slot - (slot / G) * G;
  } else {
    t = blockIdx.x / G;
    g = blockIdx.x - t * G;
  }
  if constexpr (XY)
    cluster_two_opt_tour<kClusterBlock, PP>(tours + (size_t)t * n, n, IntCoordDist{ixy}, ct, kp, k, max_iter, G, g,
                                            sync + (size_t)t * kClusterWords, stats,
                                            cost != nullptr ? cost + t : nullptr, lds, red, red64, s_best, P, sync,
                                            cost, best);
  else
    cluster_two_opt_tour<kClusterBlock, PP>(tours + (size_t)t * n, n, MatrixDist{D, ld}, ct, kp, k, max_iter, G, g,
                                            sync + (size_t)t * kClusterWords, stats,
                                            cost != nullptr ? cost + t : nullptr, lds, red, red64, s_best, P, sync,
                                            cost, best);
}

// Packed candidate table: out[u, r] = {N[u, r], D[u, N[u, r]]} for r < k, {-1, 0} for k <= r < KP.
__global__ void cand_table_kernel(const int32_t* __restrict__ cand, int n, int k, int kp,
                                  const int32_t* __restrict__ D, int ld, int2* __restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * kp) return;
  const int u = idx / kp, r = idx - u * kp;
  int2 v = make_int2(-1, 0);
  if (r < k) {
    const int nb = cand[(size_t)u * k + r];
    if (nb >= 0 && nb < n) v = make_int2(nb, D[(size_t)u * ld + nb]);
  }
  out[idx] = v;
}


// Multi-move acceptance (two_opt.py:491-563): of all improving pairs of one scan, greedily accept
// windows [i, j] in (delta, scan index) order unless they intersect an accepted window, then apply
// every accepted reversal.  The greedy is computed without materialising the improving pairs:
// the first accepted window is the scan's overall best pair; accepting [a, b] inside a free run
// [L, R] leaves the runs [L, a-1] and [b+1, R], whose greedy choices are independent of each
// other (no window can span both), so each round takes the best pair inside every live run, in
// parallel, until no run holds an improving pair.  Deltas stay those of the step's start tour
// (reversals are applied after the last round), as in the reference.
//
// LDS per position p: run[p] = start L of p's free run (-1 once covered or dead), rend[L] = end R
// of the run starting at L, key[L] = best (delta, index) inside that run; moves[] = accepted
// windows (at most n / 3: each covers >= 3 positions).
constexpr int kMultiBlock = 256;
constexpr int kMultiWaves = kMultiBlock / kWave;

__global__ __launch_bounds__(kMultiBlock) void two_opt_multi_kernel(int32_t* __restrict__ tours, int n,
                                                                    const int32_t* __restrict__ D, int ld,
                                                                    const int32_t* __restrict__ cand, int k,
                                                                    int max_iter, int64_t* __restrict__ stats) {
  extern __shared__ int4 lds_multi[];
  __shared__ int s_count, s_total;
  __shared__ int64_t red64[kMultiWaves];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  int4* s_moves = lds_multi;                                             // n / 3 + 1 windows (a, b, L, R)
  uint64_t* s_key = reinterpret_cast<uint64_t*>(s_moves + n / 3 + 1);  // n
  int32_t* s_tour = reinterpret_cast<int32_t*>(s_key + n);              // n + 1
  int32_t* s_edge = s_tour + n + 1;                                     // n
  int32_t* s_run = s_edge + n;                                          // n
  int32_t* s_rend = s_run + n;                                          // n (also the reversal copy)
  int32_t* s_pos = s_rend + n;                                          // n (candidate mode)
  int32_t* tour = tours + (size_t)blockIdx.x * n;
  for (int x = tid; x < n; x += kMultiBlock) {
    const int v = clamp_id(tour[x], n);
    s_tour[x] = v;
    if (x == 0) s_tour[n] = v;
  }
  __syncthreads();
  const int rows = n - 2;
  int64_t my_deltas = 0;
  int moves = 0, passes = 0;
  for (int it = 0; max_iter < 0 || it < max_iter; ++it) {
    ++passes;
    for (int x = tid; x < n; x += kMultiBlock) {
      const int a = s_tour[x];
      s_edge[x] = D[(size_t)a * ld + s_tour[x + 1]];
      if (cand != nullptr) s_pos[a] = x;
      s_run[x] = 0;
    }
    if (tid == 0) {
      s_rend[0] = n - 1;
      s_total = 0;
    }
    __syncthreads();
    for (int round = 0;; ++round) {
      for (int x = tid; x < n; x += kMultiBlock)
        if (s_run[x] == x) s_key[x] = kNoMove;
      if (tid == 0) s_count = 0;
      __syncthreads();
      // best improving pair inside each live run
      for (int i = wid; i < rows; i += kMultiWaves) {
        const int L = s_run[i];
        if (L < 0) continue;
        const int R = s_rend[L];
        const int lim = (i == 0) ? n - 1 : n;
        const int jmax = min(lim - 1, R);
        if (jmax < i + 2) continue;
        const int a = s_tour[i], b = s_tour[i + 1];
        const int eab = s_edge[i];
        const int32_t* __restrict__ ra = D + (size_t)a * ld;
        const int32_t* __restrict__ rb = D + (size_t)b * ld;
        uint64_t best = kNoMove;
        if (cand == nullptr) {
          const uint32_t base = (uint32_t)i * (uint32_t)n;
          for (int j = i + 2 + lane; j <= jmax; j += kWave) {
            const int delta = ra[s_tour[j]] + rb[s_tour[j + 1]] - eab - s_edge[j];
            if (delta < 0) {
              const uint64_t key = pack_key(delta, base + (uint32_t)j);
              best = key < best ? key : best;
            }
          }
          if (round == 0 && lane == 0) my_deltas += jmax - i - 1;
        } else {
          for (int r = lane; r < k; r += kWave) {
            const int nb = cand[(size_t)a * k + r];
            const int j = (nb >= 0 && nb < n) ? s_pos[nb] : -1;
            if (j > i + 1 && j <= jmax) {
              if (round == 0) ++my_deltas;
              const int delta = ra[nb] + rb[s_tour[j + 1]] - eab - s_edge[j];
              if (delta < 0) {
                const uint64_t key = pack_key(delta, (uint32_t)i * (uint32_t)k + (uint32_t)r);
                best = key < best ? key : best;
              }
            }
          }
        }
        best = wave_min_u64(best);
        if (lane == 0 && best != kNoMove) atomicMin((unsigned long long*)&s_key[L], (unsigned long long)best);
      }
      __syncthreads();
      // decode every run's choice into its window (a, b), packed as (a << 32 | b)
      for (int x = tid; x < n; x += kMultiBlock) {
        if (s_run[x] == x && s_key[x] != kNoMove) {
          const uint32_t idx = key_index(s_key[x]);
          int a, b;
          if (cand == nullptr) {
            a = (int)(idx / (uint32_t)n);
            b = (int)(idx - (uint32_t)a * (uint32_t)n);
          } else {
            a = (int)(idx / (uint32_t)k);
            b = s_pos[cand[(size_t)s_tour[a] * k + (int)(idx - (uint32_t)a * (uint32_t)k)]];
          }
          s_key[x] = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
        }
      }
      __syncthreads();
      // relabel positions; the first position of each accepted window records the move
      for (int x = tid; x < n; x += kMultiBlock) {
        const int L = s_run[x];
        if (L < 0) continue;
        const uint64_t w = s_key[L];
        if (w == kNoMove) {
          s_run[x] = -1;  // run without improving pairs: done
          continue;
        }
        const int a = (int)(w >> 32), b = (int)(uint32_t)w;
        if (x == a) {
          const int m = atomicAdd(&s_count, 1);
          s_moves[s_total + m] = make_int4(a, b, L, s_rend[L]);
        }
        s_run[x] = x < a ? L : (x <= b ? -1 : b + 1);
      }
      __syncthreads();
      const int cnt = s_count;
      if (cnt == 0) break;
      for (int m = tid; m < cnt; m += kMultiBlock) {
        const int4 mv = s_moves[s_total + m];
        if (mv.x > mv.z) s_rend[mv.z] = mv.x - 1;
        if (mv.y < mv.w) s_rend[mv.y + 1] = mv.w;
      }
      __syncthreads();
      if (tid == 0) s_total += cnt;
      __syncthreads();
    }
    const int total = s_total;
    if (total == 0) break;
    moves += total;
    // apply the disjoint reversals of positions a+1 .. b
    for (int x = tid; x < n; x += kMultiBlock) s_rend[x] = s_tour[x];
    __syncthreads();
    for (int m = 0; m < total; ++m) {
      const int a = s_moves[m].x, b = s_moves[m].y;
      for (int p = a + 1 + tid; p <= b; p += kMultiBlock) s_tour[p] = s_rend[a + 1 + b - p];
    }
    __syncthreads();
  }
  for (int x = tid; x < n; x += kMultiBlock) tour[x] = s_tour[x];
  if (stats != nullptr) {
    const int64_t all = block_sum_i64(my_deltas, red64);
    if (tid == 0) {
      atomicAdd((unsigned long long*)&stats[0], (unsigned long long)all);
      atomicAdd((unsigned long long*)&stats[1], (unsigned long long)moves);
      atomicAdd((unsigned long long*)&stats[2], (unsigned long long)passes);
    }
  }
}

size_t multi_lds_bytes(int n, bool use_cand) {
  return (size_t)n * sizeof(uint64_t) + (size_t)(4 * n + 1) * sizeof(int32_t) + (size_t)(n / 3 + 1) * sizeof(int4) +
         (use_cand ? (size_t)n * sizeof(int32_t) : 0) + 16;
}

}  // namespace tspmi

extern "C" int tspmi_two_opt(int32_t* tours, int P, int n, const int32_t* D, int ld, const int32_t* cand, int k,
                             int mode, int max_iter, int64_t* stats, void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_two_opt: negative shape P=%d n=%d", P, n);
  TSPMI_CHECK_ARG(mode == TSPMI_MODE_BEST || mode == TSPMI_MODE_FIRST || mode == TSPMI_MODE_MULTI,
                  "tspmi_two_opt: bad mode %d", mode);
  if (P == 0 || n < 4 || max_iter == 0) return TSPMI_OK;  // reference: n < 4 or no budget is a copy
  TSPMI_CHECK_ARG(tours != nullptr && D != nullptr, "tspmi_two_opt: null tours or distances");
  TSPMI_CHECK_ARG(ld >= n, "tspmi_two_opt: ld=%d < n=%d", ld, n);
  TSPMI_CHECK_ARG((long)n * ld < kMaxMatrixElems, "tspmi_two_opt: n * ld = %ld >= 2^30", (long)n * ld);
  TSPMI_CHECK_ARG((long)n * (long)n < (1L << 32), "tspmi_two_opt: n=%d too large for 32-bit scan indices", n);
  const bool use_cand = cand != nullptr && k > 0;
  if (use_cand) {
    TSPMI_CHECK_ARG(ld == n, "tspmi_two_opt: candidate mode needs ld == n (got ld=%d n=%d)", ld, n);
    TSPMI_CHECK_ARG((long)(n - 2) * k < (1L << 31), "tspmi_two_opt: (n-2)*k too large");
  }
  hipStream_t s = (hipStream_t)stream;
  if (mode == TSPMI_MODE_MULTI) {
    const size_t ml = multi_lds_bytes(n, use_cand);
    TSPMI_CHECK_ARG(ml <= (size_t)kMaxLds - 512, "tspmi_two_opt: multi_move at n=%d needs %zu B of LDS (max %d)", n, ml,
                    kMaxLds);
    hipLaunchKernelGGL(two_opt_multi_kernel, dim3(P), dim3(kMultiBlock), ml, s, tours, n, D, ld, use_cand ? cand : nullptr,
                       use_cand ? k : 0, max_iter, stats);
    TSPMI_CHECK_LAUNCH("tspmi_two_opt launch");
    return TSPMI_OK;
  }
  // The diagonal scan reads up to QMAX * 256 + 64 entries past the tour/edge arrays (zero padding).
  constexpr int kQ = 8;
  const int q = (!use_cand && mode == TSPMI_MODE_BEST) ? (n - 2 <= kTwoOptWaves * 4 * kWave ? 4
                                                          : (n - 2 <= kTwoOptWaves * kQ * kWave ? kQ : 0))
                                                       : 0;
  const int pad = q > 0 ? q * kTwoOptBlock + kWave : 0;
  const size_t lds = ((size_t)(use_cand ? 3 : 2) * n + 1 + 2 * (size_t)pad) * sizeof(int32_t);
  TSPMI_CHECK_ARG(lds + sizeof(TwoOptShared) <= (size_t)kMaxLds,
                  "tspmi_two_opt: n=%d needs %zu B of LDS (max %d)", n, lds, kMaxLds);
  const size_t staged_lds = (size_t)(3 * n + 1) * sizeof(int32_t);
  if (!use_cand && mode == TSPMI_MODE_BEST && q == 0 && n - 2 <= 10 * kWideBlock + 1 &&
      staged_lds + 1024 <= (size_t)kMaxLds) {
    if (n - 2 <= 6 * kWideBlock + 1)
      hipLaunchKernelGGL(two_opt_staged_kernel<6>, dim3(P), dim3(kWideBlock), staged_lds, s, tours, n, D, ld, max_iter,
                         stats);
    else
      hipLaunchKernelGGL(two_opt_staged_kernel<10>, dim3(P), dim3(kWideBlock), staged_lds, s, tours, n, D, ld, max_iter,
                         stats);
  } else if (!use_cand && mode == TSPMI_MODE_BEST && q == 0 && n - 2 <= kWideWaves * 16 * kWave &&
      (size_t)(2 * n + 1) * sizeof(int32_t) + 1024 <= (size_t)kMaxLds) {
    const size_t wl = (size_t)(2 * n + 1) * sizeof(int32_t);
    if (n - 2 <= kWideWaves * 10 * kWave)
      hipLaunchKernelGGL(two_opt_wide_kernel<10>, dim3(P), dim3(kWideBlock), wl, s, tours, n, D, ld, max_iter, stats);
    else
      hipLaunchKernelGGL(two_opt_wide_kernel<16>, dim3(P), dim3(kWideBlock), wl, s, tours, n, D, ld, max_iter, stats);
  } else if (q == 4) {
    hipLaunchKernelGGL(two_opt_kernel<4>, dim3(P), dim3(kTwoOptBlock), lds, s, tours, n, D, ld, nullptr, 0, mode,
                       max_iter, pad, stats);
  } else if (q == kQ) {
    hipLaunchKernelGGL(two_opt_kernel<kQ>, dim3(P), dim3(kTwoOptBlock), lds, s, tours, n, D, ld, nullptr, 0, mode,
                       max_iter, pad, stats);
  } else {
    hipLaunchKernelGGL(two_opt_kernel<0>, dim3(P), dim3(kTwoOptBlock), lds, s, tours, n, D, ld,
                       use_cand ? cand : nullptr, use_cand ? k : 0, mode, max_iter, 0, stats);
  }
  TSPMI_CHECK_LAUNCH("tspmi_two_opt launch");
  return TSPMI_OK;
}

namespace {
// Row width of the packed candidate table for k candidates (8, 16 or 32).
int cand_lanes(int k) { return k <= 8 ? 8 : (k <= 16 ? 16 : 32); }
}  // namespace

namespace tspmi {

// ---- Node labels in Hilbert-curve order ------------------------------------------------------
//
// The gathered rows of the full scan read D[t_{i+2}][t_{j+2}] for 64 consecutive positions j at
// once; with the nodes numbered along a space-filling curve of their coordinates, the tour
// neighbours of a partly optimised tour have nearby numbers, so a wave's 64 columns fall into fewer
// cache lines and each gather costs the texture addresser fewer cycles.  Renumbering changes no
// delta and no position, so the scan makes the same moves.

constexpr int kOrderBlock = 1024;
constexpr int kOrderMaxN = 8192;  // keys held in LDS by spatial_order_kernel

// One workgroup: the coordinates' bounding box, each node's Hilbert index on a 2^16 grid over the
// box (one scale for both axes), then order[r] = the node of rank r and rank[u] = the rank of u by
// (index, node), a stable order.
__global__ __launch_bounds__(kOrderBlock) void spatial_order_kernel(const double2* __restrict__ xy, int n,
                                                                   int32_t* __restrict__ order,
                                                                   int32_t* __restrict__ rank) {
  __shared__ uint32_t s_key[kOrderMaxN];
  __shared__ double s_box[4][kOrderBlock / kWave];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  double lo_x = __builtin_inf(), lo_y = __builtin_inf(), hi_x = -__builtin_inf(), hi_y = -__builtin_inf();
  for (int u = tid; u < n; u += kOrderBlock) {
    const double2 p = xy[u];
    lo_x = fmin(lo_x, p.x);
    lo_y = fmin(lo_y, p.y);
    hi_x = fmax(hi_x, p.x);
    hi_y = fmax(hi_y, p.y);
  }
  for (int m = 32; m >= 1; m >>= 1) {
    lo_x = fmin(lo_x, __shfl_xor(lo_x, m, kWave));
    lo_y = fmin(lo_y, __shfl_xor(lo_y, m, kWave));
    hi_x = fmax(hi_x, __shfl_xor(hi_x, m, kWave));
    hi_y = fmax(hi_y, __shfl_xor(hi_y, m, kWave));
  }
  if (lane == 0) {
    s_box[0][wid] = lo_x;
    s_box[1][wid] = lo_y;
    s_box[2][wid] = hi_x;
    s_box[3][wid] = hi_y;
  }
  __syncthreads();
  for (int w = 0; w < kOrderBlock / kWave; ++w) {
    lo_x = fmin(lo_x, s_box[0][w]);
    lo_y = fmin(lo_y, s_box[1][w]);
    hi_x = fmax(hi_x, s_box[2][w]);
    hi_y = fmax(hi_y, s_box[3][w]);
  }
  const double span = fmax(fmax(hi_x - lo_x, hi_y - lo_y), 1e-300);
  const double scale = 65535.0 / span;
  for (int u = tid; u < n; u += kOrderBlock) {
    const double2 p = xy[u];
    // NaN or infinite coordinates (an invalid input) land on cell 0 rather than anywhere undefined
    const double fx = (p.x - lo_x) * scale, fy = (p.y - lo_y) * scale;
    const uint32_t qx = fx >= 0.0 && fx <= 65535.0 ? (uint32_t)fx : 0u;
    const uint32_t qy = fy >= 0.0 && fy <= 65535.0 ? (uint32_t)fy : 0u;
    s_key[u] = hilbert_index16(qx, qy);
  }
  __syncthreads();
  for (int u = tid; u < n; u += kOrderBlock) {
    const uint32_t ku = s_key[u];
    int r = 0;
    for (int v = 0; v < n; ++v) {
      const uint32_t kv = s_key[v];
      r += (kv < ku || (kv == ku && v < u)) ? 1 : 0;
    }
    rank[u] = r;
    order[r] = u;
  }
}

// D2[r][c] = D[order[r]][order[c]] (row stride n) and xy2[r] = xy[order[r]].
__global__ void permute_instance_kernel(const int32_t* __restrict__ D, int ld, const double2* __restrict__ xy, int n,
                                        const int32_t* __restrict__ order, int32_t* __restrict__ D2,
                                        double2* __restrict__ xy2) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx - (int64_t)r * n);
  const int a = order[r], b = order[c];
  D2[idx] = D[(size_t)a * ld + b];
  if (r == 0) xy2[c] = xy[b];
}

// tours[x] = map[tours[x]] (ids outside [0, n) are read as node 0, as the scan itself does).
__global__ void relabel_kernel(int32_t* __restrict__ tours, int64_t count, int n, const int32_t* __restrict__ map) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < count) tours[idx] = map[clamp_id(tours[idx], n)];
}

// Workspace layout of tspmi_two_opt_full_euc2d_ordered: D2 (n * n int32), xy2 (n double2), order
// and rank (n int32 each), each part 256-byte aligned.
__host__ __device__ inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
inline size_t ordered_ws_bytes(int n) {
  return align256((size_t)n * n * 4) + align256((size_t)n * 16) + 2 * align256((size_t)n * 4);
}

}  // namespace tspmi

extern "C" int tspmi_two_opt_full_euc2d(int32_t* tours, int P, int n, const int32_t* D, int ld, const double* xy,
                                   int max_iter, int64_t* stats, void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_two_opt_full_euc2d: negative shape P=%d n=%d", P, n);
  if (P == 0 || n < 4 || max_iter == 0) return TSPMI_OK;  // reference: n < 4 or no budget is a copy
  TSPMI_CHECK_ARG(tours != nullptr && D != nullptr && xy != nullptr, "tspmi_two_opt_full_euc2d: null pointer");
  TSPMI_CHECK_ARG(((uintptr_t)xy & 15) == 0, "tspmi_two_opt_full_euc2d: xy must be 16-byte aligned");
  TSPMI_CHECK_ARG(ld >= n && (long)n * ld < kMaxMatrixElems, "tspmi_two_opt_full_euc2d: bad ld=%d for n=%d", ld, n);
  TSPMI_CHECK_ARG(n - 2 <= kTwoOptWaves * 8 * kWave, "tspmi_two_opt_full_euc2d: n=%d above 2050", n);
  const int q = n - 2 <= kTwoOptWaves * 4 * kWave ? 4 : 8;
  const int pad = q * kTwoOptBlock + kWave;
  const size_t lds = ((size_t)mix_coord_words(n, pad) + 2 * n + 1 + 2 * (size_t)pad) * sizeof(int32_t);
  TSPMI_CHECK_ARG(lds + 1024 <= (size_t)kMaxLds, "tspmi_two_opt_full_euc2d: n=%d needs %zu B of LDS", n, lds);
  hipStream_t s = (hipStream_t)stream;
  const double2* p = reinterpret_cast<const double2*>(xy);
  if (q == 4)
    hipLaunchKernelGGL(two_opt_mix_kernel<4>, dim3(P), dim3(kTwoOptBlock), lds, s, tours, n, D, ld, p, max_iter, pad,
                       stats);
  else
    hipLaunchKernelGGL(two_opt_mix_kernel<8>, dim3(P), dim3(kTwoOptBlock), lds, s, tours, n, D, ld, p, max_iter, pad,
                       stats);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_full_euc2d launch");
  return TSPMI_OK;
}

namespace tspmi {

// One workgroup: keys[u] = the Hilbert index of int32 coordinates xy[u] on a 2^16 grid over their
// bounding box (one scale for both axes), as a double (exact below 2^53).
__global__ __launch_bounds__(kOrderBlock) void hilbert_keys_i32_kernel(const int2* __restrict__ xy, int n,
                                                                      double* __restrict__ keys) {
  __shared__ int s_box[4];
  const int tid = threadIdx.x;
  if (tid < 4) s_box[tid] = 0x7fffffff;
  __syncthreads();
  int lo_x = 0x7fffffff, lo_y = 0x7fffffff, nhi_x = 0x7fffffff, nhi_y = 0x7fffffff;  // hi as -max
  for (int u = tid; u < n; u += kOrderBlock) {
    const int2 p = xy[u];
    lo_x = min(lo_x, p.x);
    lo_y = min(lo_y, p.y);
    nhi_x = min(nhi_x, -p.x);
    nhi_y = min(nhi_y, -p.y);
  }
  atomicMin(&s_box[0], lo_x);
  atomicMin(&s_box[1], lo_y);
  atomicMin(&s_box[2], nhi_x);
  atomicMin(&s_box[3], nhi_y);
  __syncthreads();
  const double span = fmax(fmax((double)-s_box[2] - s_box[0], (double)-s_box[3] - s_box[1]), 1.0);
  const double scale = 65535.0 / span;
  for (int u = tid; u < n; u += kOrderBlock) {
    const int2 p = xy[u];
    const uint32_t qx = (uint32_t)(((double)p.x - s_box[0]) * scale);
    const uint32_t qy = (uint32_t)(((double)p.y - s_box[1]) * scale);
    keys[u] = (double)hilbert_index16(qx, qy);
  }
}

// rank[order[r]] = r.
__global__ void invert_order_kernel(const int32_t* __restrict__ order, int n, int32_t* __restrict__ rank) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) rank[order[r]] = r;
}

// The renumbered candidate table and coordinates: row r is node order[r], its neighbours renamed by
// rank (the padding entries, neighbour -1, kept); xy2[r] = xy[order[r]].
__global__ void permute_table_kernel(const int2* __restrict__ table, int kp, int n, const int32_t* __restrict__ order,
                                     const int32_t* __restrict__ rank, const int2* __restrict__ xy,
                                     int2* __restrict__ table2, int2* __restrict__ xy2) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * kp) return;
  const int r = idx / kp, q = idx - r * kp;
  const int u = order[r];
  int2 e = table[(size_t)u * kp + q];
  e.x = (unsigned)e.x < (unsigned)n ? rank[e.x] : e.x;
  table2[idx] = e;
  if (q == 0) xy2[r] = xy[u];
}

// Workspace of tspmi_two_opt_packed_euc2d_i32_ordered: keys (n double), order and rank (n int32
// each), coordinates (n int2) and table (n * kp int2), each part 256-byte aligned.
inline size_t packed_ordered_ws_bytes(int n, int kp) {
  return align256((size_t)n * 8) + 2 * align256((size_t)n * 4) + align256((size_t)n * 8) +
         align256((size_t)n * kp * 8);
}

}  // namespace tspmi

extern "C" size_t tspmi_two_opt_full_euc2d_ws_bytes(int n) { return n > 0 ? tspmi::ordered_ws_bytes(n) : 0; }

extern "C" size_t tspmi_two_opt_packed_euc2d_ws_bytes(int n, int k) {
  return n > 0 && k >= 1 && k <= 32 ? tspmi::packed_ordered_ws_bytes(n, cand_lanes(k)) : 0;
}

extern "C" int tspmi_two_opt_packed_euc2d_i32_ordered(int32_t* tours, int P, int n, const int32_t* ixy,
                                                      const int32_t* table, int k, int max_iter, int64_t* stats,
                                                      double* cost, int64_t* pass_log, int log_moves, void* ws,
                                                      size_t ws_bytes, void* stream) {
  using namespace tspmi;
  // Shapes the renumbering does not apply to (and every argument error) go to the plain call as
  // given, and so do tours whose node records fit LDS: that scan reads no coordinates from global
  // memory, and renumbering measured no gain there (n = 1000 to 5000); above, at n = 10000, P = 256,
  // k = 8 / 20, it saves 9 / 10 %.
  if (P <= 0 || n < 4 || max_iter == 0 || k < 1 || k > 32 || ws == nullptr || tours == nullptr || ixy == nullptr ||
      table == nullptr || (size_t)rec_lds_words(n) * sizeof(int32_t) + 1024 <= (size_t)kMaxLds)
    return tspmi_two_opt_packed_euc2d_i32(tours, P, n, ixy, table, k, max_iter, stats, cost, pass_log, log_moves,
                                          stream);
  const int kp = cand_lanes(k);
  TSPMI_CHECK_ARG(((uintptr_t)ws & 255) == 0, "tspmi_two_opt_packed_euc2d_i32_ordered: ws must be 256-byte aligned");
  TSPMI_CHECK_ARG(ws_bytes >= packed_ordered_ws_bytes(n, kp),
                  "tspmi_two_opt_packed_euc2d_i32_ordered: workspace of %zu B, %zu needed", ws_bytes,
                  packed_ordered_ws_bytes(n, kp));
  TSPMI_CHECK_ARG(((uintptr_t)ixy & 7) == 0 && ((uintptr_t)table & 7) == 0,
                  "tspmi_two_opt_packed_euc2d_i32_ordered: xy and table must be 8-byte aligned");
  TSPMI_CHECK_ARG((size_t)(3 * n + 1) * sizeof(int32_t) + 1024 <= (size_t)kMaxLds,
                  "tspmi_two_opt_packed_euc2d_i32_ordered: n=%d exceeds the LDS limit", n);
  TSPMI_CHECK_ARG(pass_log == nullptr || log_moves >= 0, "tspmi_two_opt_packed_euc2d_i32_ordered: bad pass log");
  char* w = static_cast<char*>(ws);
  double* keys = reinterpret_cast<double*>(w);
  int32_t* order = reinterpret_cast<int32_t*>(w + align256((size_t)n * 8));
  int32_t* rank = reinterpret_cast<int32_t*>(w + align256((size_t)n * 8) + align256((size_t)n * 4));
  int2* xy2 = reinterpret_cast<int2*>(w + align256((size_t)n * 8) + 2 * align256((size_t)n * 4));
  int2* table2 = reinterpret_cast<int2*>(w + align256((size_t)n * 8) + 2 * align256((size_t)n * 4) +
                                         align256((size_t)n * 8));
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(hilbert_keys_i32_kernel, dim3(1), dim3(kOrderBlock), 0, s, reinterpret_cast<const int2*>(ixy), n,
                     keys);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_packed_euc2d_i32_ordered keys launch");
  int rc = tspmi_rank_select(keys, n, n, order, stream);
  if (rc != TSPMI_OK) return rc;
  hipLaunchKernelGGL(invert_order_kernel, dim3((n + 255) / 256), dim3(256), 0, s, order, n, rank);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_packed_euc2d_i32_ordered rank launch");
  hipLaunchKernelGGL(permute_table_kernel, dim3((n * kp + 255) / 256), dim3(256), 0, s,
                     reinterpret_cast<const int2*>(table), kp, n, order, rank, reinterpret_cast<const int2*>(ixy),
                     table2, xy2);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_packed_euc2d_i32_ordered permute launch");
  const int64_t count = (int64_t)P * n;
  const dim3 grid((unsigned)((count + 255) / 256));
  hipLaunchKernelGGL(relabel_kernel, grid, dim3(256), 0, s, tours, count, n, rank);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_packed_euc2d_i32_ordered relabel launch");
  rc = tspmi_two_opt_packed_euc2d_i32(tours, P, n, reinterpret_cast<const int32_t*>(xy2),
                                      reinterpret_cast<const int32_t*>(table2), k, max_iter, stats, cost, pass_log,
                                      log_moves, stream);
  if (rc != TSPMI_OK) return rc;
  hipLaunchKernelGGL(relabel_kernel, grid, dim3(256), 0, s, tours, count, n, order);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_packed_euc2d_i32_ordered relabel launch");
  return TSPMI_OK;
}

extern "C" int tspmi_two_opt_full_euc2d_ordered(int32_t* tours, int P, int n, const int32_t* D, int ld,
                                                const double* xy, int max_iter, int64_t* stats, void* ws,
                                                size_t ws_bytes, void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_two_opt_full_euc2d_ordered: negative shape P=%d n=%d", P, n);
  if (P == 0 || n < 4 || max_iter == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(tours != nullptr && D != nullptr && xy != nullptr && ws != nullptr,
                  "tspmi_two_opt_full_euc2d_ordered: null pointer");
  TSPMI_CHECK_ARG(((uintptr_t)xy & 15) == 0 && ((uintptr_t)ws & 255) == 0,
                  "tspmi_two_opt_full_euc2d_ordered: xy must be 16-byte and ws 256-byte aligned");
  TSPMI_CHECK_ARG(ld >= n && (long)n * ld < kMaxMatrixElems, "tspmi_two_opt_full_euc2d_ordered: bad ld=%d for n=%d",
                  ld, n);
  TSPMI_CHECK_ARG(n <= kOrderMaxN && n - 2 <= kTwoOptWaves * 8 * kWave,
                  "tspmi_two_opt_full_euc2d_ordered: n=%d above 2050", n);
  TSPMI_CHECK_ARG(ws_bytes >= ordered_ws_bytes(n), "tspmi_two_opt_full_euc2d_ordered: workspace of %zu B, %zu needed",
                  ws_bytes, ordered_ws_bytes(n));
  {  // the scan's own LDS check, made before the tours are renumbered in place
    const int q = n - 2 <= kTwoOptWaves * 4 * kWave ? 4 : 8;
    const int pad = q * kTwoOptBlock + kWave;
    const size_t lds = ((size_t)mix_coord_words(n, pad) + 2 * n + 1 + 2 * (size_t)pad) * sizeof(int32_t);
    TSPMI_CHECK_ARG(lds + 1024 <= (size_t)kMaxLds, "tspmi_two_opt_full_euc2d_ordered: n=%d needs %zu B of LDS", n, lds);
  }
  char* w = static_cast<char*>(ws);
  int32_t* D2 = reinterpret_cast<int32_t*>(w);
  double2* xy2 = reinterpret_cast<double2*>(w + align256((size_t)n * n * 4));
  int32_t* order = reinterpret_cast<int32_t*>(w + align256((size_t)n * n * 4) + align256((size_t)n * 16));
  int32_t* rank = order + align256((size_t)n * 4) / 4;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(spatial_order_kernel, dim3(1), dim3(kOrderBlock), 0, s, reinterpret_cast<const double2*>(xy), n,
                     order, rank);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_full_euc2d_ordered order launch");
  const int64_t cells = (int64_t)n * n;
  hipLaunchKernelGGL(permute_instance_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, D, ld,
                     reinterpret_cast<const double2*>(xy), n, order, D2, xy2);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_full_euc2d_ordered permute launch");
  const int64_t count = (int64_t)P * n;
  const dim3 grid((unsigned)((count + 255) / 256));
  hipLaunchKernelGGL(relabel_kernel, grid, dim3(256), 0, s, tours, count, n, rank);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_full_euc2d_ordered relabel launch");
  const int rc = tspmi_two_opt_full_euc2d(tours, P, n, D2, n, reinterpret_cast<const double*>(xy2), max_iter, stats,
                                          stream);
  if (rc != TSPMI_OK) return rc;
  hipLaunchKernelGGL(relabel_kernel, grid, dim3(256), 0, s, tours, count, n, order);
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_full_euc2d_ordered relabel launch");
  return TSPMI_OK;
}


extern "C" int tspmi_candidate_table(const int32_t* cand, int n, int k, const int32_t* D, int ld, int32_t* out,
                                     void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(n >= 0 && k >= 1 && k <= 32 && ld >= n, "tspmi_candidate_table: bad shape n=%d k=%d ld=%d", n, k,
                  ld);
  if (n == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(cand && D && out, "tspmi_candidate_table: null pointer");
  const int kp = cand_lanes(k);
  const int total = n * kp;
  hipLaunchKernelGGL(cand_table_kernel, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, cand, n, k, kp,
                     D, ld, reinterpret_cast<int2*>(out));
  TSPMI_CHECK_LAUNCH("tspmi_candidate_table launch");
  return TSPMI_OK;
}

extern "C" int tspmi_two_opt_packed_cost(int32_t* tours, int P, int n, const int32_t* D, int ld,
                                         const int32_t* table, int k, int max_iter, int64_t* stats, double* cost,
                                         void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_two_opt_packed: negative shape");
  TSPMI_CHECK_ARG(k >= 1 && k <= 32, "tspmi_two_opt_packed: k=%d outside [1, 32]", k);
  if (P == 0) return TSPMI_OK;
  if (n < 4 || max_iter == 0) {
    // Nothing moves; the costs are still asked for.
    return cost != nullptr ? tspmi_tour_cost(tours, P, n, D, ld, 1, cost, stream) : TSPMI_OK;
  }
  TSPMI_CHECK_ARG(tours && D && table, "tspmi_two_opt_packed: null pointer");
  TSPMI_CHECK_ARG(ld >= n && (long)(n - 2) * k < (1L << 31) && (long)n * ld < kMaxMatrixElems,
                  "tspmi_two_opt_packed: bad ld/n/k (n * ld must stay below 2^30)");
  TSPMI_CHECK_ARG(((uintptr_t)table & 7) == 0, "tspmi_two_opt_packed: table must be 8-byte aligned");
  const size_t lds = (size_t)(3 * n + 1) * sizeof(int32_t);
  TSPMI_CHECK_ARG(lds + 1024 <= (size_t)kMaxLds, "tspmi_two_opt_packed: n=%d exceeds the LDS limit", n);
  hipStream_t s = (hipStream_t)stream;
  const int2* ct = reinterpret_cast<const int2*>(table);
  const int kp = cand_lanes(k);
  // Size the workgroup so that the tours fill the compute units: one tour per CU gets 16 waves,
  // up to two per CU get 8 waves each, more get 4.
  const int cus = device_cu_count();
  const int block = P <= cus ? 1024 : (P <= 2 * cus ? 512 : 256);
  // One tour per CU and at most 20 pairs per thread: the register-resident form.
  const int pp = (n * k + 1023) / 1024;
  const size_t reg_lds = (size_t)(5 * n + 1) * sizeof(int32_t);
  if (block == 1024 && pp <= 20 && n <= (1 << kRegJBits) && reg_lds + 1024 <= (size_t)kMaxLds) {
#define TSPMI_LAUNCH_REG(PPV)                                                                                       \
  hipLaunchKernelGGL((two_opt_reg_kernel<PPV>), dim3(P), dim3(1024), reg_lds, s, tours, n, D, ld, ct, kp, k, max_iter, \
                     stats, cost)
    if (pp <= 4) TSPMI_LAUNCH_REG(4);
    else if (pp <= 8) TSPMI_LAUNCH_REG(8);
    else if (pp <= 12) TSPMI_LAUNCH_REG(12);
    else if (pp <= 16) TSPMI_LAUNCH_REG(16);
    else TSPMI_LAUNCH_REG(20);
#undef TSPMI_LAUNCH_REG
    TSPMI_CHECK_LAUNCH("tspmi_two_opt_packed launch");
    return TSPMI_OK;
  }
#define TSPMI_LAUNCH_CAND(BL, KPV, KLV)                                                                          \
  hipLaunchKernelGGL((two_opt_cand_kernel<BL, KPV, 4, KLV>), dim3(P), dim3(BL), lds, s, tours, n, D, ld, ct, k,       \
                     max_iter, stats, cost)
  // (A 16-way unroll at 1024 threads was measured slower on 16-tour elite calls: 16 vs 14.5 us per pass.)
  // (Holding the table in LDS for these launches saved 7 % per pass but cost 5.6 us per launch to
  // fill, measured on 16-tour elite calls; not kept.  A node-indexed streaming scan -- two 16-byte
  // node records per pair, no KP padding, a 64-bit key minimum per pair -- ran 1.36e11 deltas/s
  // against this kernel's 2.0e11 at P = 1024, n = 1000, k = 20; not kept either.)
  if (block == 1024) {
    TSPMI_CAND_BY_K(TSPMI_LAUNCH_CAND, 1024, k)
  } else if (block == 512) {
    TSPMI_CAND_BY_K(TSPMI_LAUNCH_CAND, 512, k)
  } else {
    TSPMI_CAND_BY_K(TSPMI_LAUNCH_CAND, 256, k)
  }
#undef TSPMI_LAUNCH_CAND
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_packed launch");
  return TSPMI_OK;
}

extern "C" int tspmi_two_opt_packed(int32_t* tours, int P, int n, const int32_t* D, int ld, const int32_t* table,
                                    int k, int max_iter, int64_t* stats, void* stream) {
  return tspmi_two_opt_packed_cost(tours, P, n, D, ld, table, k, max_iter, stats, nullptr, stream);
}

namespace tspmi {

// Zeroes `words` 64-bit words (a kernel rather than hipMemsetAsync, so a captured graph replays it
// correctly).
__global__ void zero_words_kernel(uint64_t* __restrict__ w, int words) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) w[i] = 0ull;
}

// failed[0] = number of tours whose cluster barrier timed out (error word of each tour's workspace).
__global__ void count_elite_errors_kernel(const uint64_t* __restrict__ sync, int P, int32_t* __restrict__ failed) {
  int c = 0;
  for (int t = threadIdx.x; t < P; t += blockDim.x)
    c += sync[(size_t)t * kClusterWords + 2 * kClusterMaxG * kClusterLine] != 0ull ? 1 : 0;
  for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, kWave);
  if (threadIdx.x == 0) failed[0] = c;
}

// The elite search of tspmi_two_opt_elite on workspace `sync`.  With `fresh`, the workspace of the
// P tours is zeroed first (so nothing of an earlier call, finished or timed out, carries over),
// and `failed` (nullable) receives the count of tours whose barrier timed out.
// With ixy (int32 coordinates, each axis spanning at most kIntCoordSpan) the distances come from
// the coordinates and D is not read.
int elite_run(int32_t* tours, int P, int n, const int32_t* D, int ld, const int32_t* ixy, const int32_t* table, int k,
              int max_iter, int64_t* stats, double* cost, int32_t* best, uint64_t* sync, bool fresh,
              int32_t* failed, hipStream_t s) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_two_opt_elite: negative shape");
  TSPMI_CHECK_ARG(k >= 1 && k <= 32, "tspmi_two_opt_elite: k=%d outside [1, 32]", k);
  const int cus = device_cu_count();
  // Largest cluster (up to 16 workgroups) that keeps one workgroup per CU over all tours, if its
  // share of pairs fits 8 registers per thread.
  int G = P > 0 ? cus / P : 0;
  G = G > 16 ? 16 : G;
  const long pairs = (long)n * k;
  const long per = G > 0 ? (pairs + (long)G * kClusterBlock - 1) / ((long)G * kClusterBlock) : 1L << 30;
  const size_t lds = (size_t)(ixy != nullptr ? 7 * n + 2 : 5 * n + 1) * sizeof(int32_t);
  const bool cluster = P > 0 && n >= 4 && max_iter != 0 && G >= 2 && per <= 8 && n <= (1 << kRegJBits) &&
                       pairs < (1L << (32 - kRegJBits)) && lds + 1024 <= (size_t)kMaxLds;
  TSPMI_CHECK_ARG(best == nullptr || cost != nullptr, "tspmi_two_opt_elite: best needs cost");
  if (!cluster) {
    int rc = ixy != nullptr
                 ? tspmi_two_opt_packed_euc2d_i32(tours, P, n, ixy, table, k, max_iter, stats, cost, nullptr, 0, s)
                 : tspmi_two_opt_packed_cost(tours, P, n, D, ld, table, k, max_iter, stats, cost, s);
    if (rc == TSPMI_OK && best != nullptr && P > 0) rc = tspmi_rank_select(cost, P, 1, best, s);
    if (rc == TSPMI_OK && fresh && failed != nullptr) {  // no barrier ran: writes failed[0] = 0
      hipLaunchKernelGGL(count_elite_errors_kernel, dim3(1), dim3(kWave), 0, s, sync, 0, failed);
      TSPMI_CHECK_LAUNCH("tspmi_two_opt_elite error count launch");
    }
    return rc;
  }
  TSPMI_CHECK_ARG(tours && (D || ixy) && table && sync, "tspmi_two_opt_elite: null pointer");
  TSPMI_CHECK_ARG(ixy != nullptr || (ld >= n && (long)n * ld < kMaxMatrixElems), "tspmi_two_opt_elite: bad ld=%d for n=%d",
                  ld, n);
  TSPMI_CHECK_ARG(((uintptr_t)table & 7) == 0 && ((uintptr_t)sync & 7) == 0 && ((uintptr_t)ixy & 7) == 0,
                  "tspmi_two_opt_elite: table, sync and coordinates must be 8-byte aligned");
  const int2* ct = reinterpret_cast<const int2*>(table);
  const int kp = cand_lanes(k);
  const int xcds = (P % kXcds == 0 && cus % kXcds == 0) ? kXcds : 1;
  if (fresh) {
    const int words = P * kClusterWords;
    hipLaunchKernelGGL(zero_words_kernel, dim3((words + 255) / 256), dim3(256), 0, s, sync, words);
    TSPMI_CHECK_LAUNCH("tspmi_two_opt_elite workspace clear launch");
  }
  const int2* cxy = reinterpret_cast<const int2*>(ixy);
#define TSPMI_LAUNCH_CLUSTER(PPV)                                                                                   \
  if (ixy != nullptr)                                                                                              \
    hipLaunchKernelGGL((two_opt_cluster_kernel<PPV, true>), dim3(P * G), dim3(kClusterBlock), lds, s, tours, n, D, ld, \
                       cxy, ct, kp, k, max_iter, G, sync, stats, cost, P, best, xcds);                             \
  else                                                                                                             \
    hipLaunchKernelGGL((two_opt_cluster_kernel<PPV>), dim3(P * G), dim3(kClusterBlock), lds, s, tours, n, D, ld, cxy, \
                       ct, kp, k, max_iter, G, sync, stats, cost, P, best, xcds)
  if (per <= 2) TSPMI_LAUNCH_CLUSTER(2);
  else if (per <= 4) TSPMI_LAUNCH_CLUSTER(4);
  else TSPMI_LAUNCH_CLUSTER(8);
#undef TSPMI_LAUNCH_CLUSTER
  TSPMI_CHECK_LAUNCH("tspmi_two_opt_elite launch");
  if (fresh && failed != nullptr) {
    hipLaunchKernelGGL(count_elite_errors_kernel, dim3(1), dim3(kWave), 0, s, sync, P, failed);
    TSPMI_CHECK_LAUNCH("tspmi_two_opt_elite error count launch");
  }
  return TSPMI_OK;
}

// The elite workspace of tspmi_two_opt_elite_noalloc: kEliteStaticTours tours of kClusterWords
// words in device memory that the library holds for the process's lifetime (540,672 bytes per GPU).
constexpr int kEliteStaticTours = TSPMI_ELITE_STATIC_TOURS;
static __device__ __align__(256) uint64_t scratch[kEliteStaticTours * kClusterWords];

// Device address of `scratch` on the current device, looked up once per device.
int elite_scratch(uint64_t** out) {
  static std::mutex mu;
  static uint64_t* addr[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "tspmi_two_opt_elite_noalloc: hipGetDevice");
  TSPMI_CHECK_ARG(dev >= 0 && dev < 64, "tspmi_two_opt_elite_noalloc: device %d outside [0, 64)", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "tspmi_two_opt_elite_noalloc: hipGetSymbolAddress(scratch)");
    addr[dev] = static_cast<uint64_t*>(p);
  }
  *out = addr[dev];
  return TSPMI_OK;
}

}  // namespace tspmi

extern "C" int tspmi_two_opt_elite(int32_t* tours, int P, int n, const int32_t* D, int ld, const int32_t* table, int k,
                                   int max_iter, int64_t* stats, double* cost, int32_t* best, uint64_t* sync,
                                   void* stream) {
  return tspmi::elite_run(tours, P, n, D, ld, nullptr, table, k, max_iter, stats, cost, best, sync, false, nullptr,
                          (hipStream_t)stream);
}

extern "C" int tspmi_two_opt_elite_xy(int32_t* tours, int P, int n, const int32_t* ixy, const int32_t* table, int k,
                                      int max_iter, int64_t* stats, double* cost, int32_t* best, uint64_t* sync,
                                      void* stream) {
  TSPMI_CHECK_ARG(ixy != nullptr || P == 0, "tspmi_two_opt_elite_xy: null coordinates");
  return tspmi::elite_run(tours, P, n, nullptr, n, ixy, table, k, max_iter, stats, cost, best, sync, false, nullptr,
                          (hipStream_t)stream);
}

extern "C" int tspmi_two_opt_elite_noalloc(int32_t* tours, int P, int n, const int32_t* D, int ld,
                                           const int32_t* table, int k, int max_iter, int64_t* stats, double* cost,
                                           int32_t* best, int32_t* failed, void* stream) {
  using namespace tspmi;
  TSPMI_CHECK_ARG(P >= 0 && P <= kEliteStaticTours,
                  "tspmi_two_opt_elite_noalloc: P=%d outside [0, %d], the static workspace's capacity", P,
                  kEliteStaticTours);
  uint64_t* ws = nullptr;
  const int rc = elite_scratch(&ws);
  if (rc != TSPMI_OK) return rc;
  return elite_run(tours, P, n, D, ld, nullptr, table, k, max_iter, stats, cost, best, ws, true, failed,
                   (hipStream_t)stream);
}
