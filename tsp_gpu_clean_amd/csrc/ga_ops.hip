// GA operators on device: tournament, PMX, inversion, the fused generation, stable rank
// selection and row gather/scatter (gfx950).
//
// Random numbers are never drawn here: the host draws them from the run's generator in the
// reference order and passes the buffers in, so every kernel is a pure function of its inputs.
#include "ga_dev.hpp"

namespace tspmi {

// One wave per output row.
__global__ __launch_bounds__(kRowBlock) void tournament_kernel(const int32_t* __restrict__ pop, int P, int n,
                                                               const double* __restrict__ fitness,
                                                               const int32_t* __restrict__ cand, int k, int rows,
                                                               int32_t* __restrict__ out,
                                                               int32_t* __restrict__ winners_out) {
  const int r = blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= rows) return;
  const int w = tournament_winner(cand + (size_t)r * k, k, fitness, P);
  if (winners_out != nullptr && lane == 0) winners_out[r] = w;
  const int32_t* src = pop + (size_t)w * n;
  int32_t* dst = out + (size_t)r * n;
  for (int x = lane; x < n; x += kWave) dst[x] = src[x];
}

// One workgroup per pair (2q, 2q+1); LDS: p1, p2, inv1, inv2 (4n ints) and, when `with_jump`,
// the PMX jump table (n more ints).
__global__ __launch_bounds__(kRowBlock) void pmx_kernel(const int32_t* __restrict__ parents, int P, int n,
                                                        const int32_t* __restrict__ cut_lo,
                                                        const int32_t* __restrict__ cut_hi,
                                                        int32_t* __restrict__ out, int with_jump) {
  extern __shared__ int32_t lds[];
  const int q = blockIdx.x;
  const int r0 = 2 * q;
  if (r0 + 1 >= P) {  // odd last row is copied (pmx_crossover.py:168-169)
    for (int x = threadIdx.x; x < n; x += blockDim.x) out[(size_t)r0 * n + x] = parents[(size_t)r0 * n + x];
    return;
  }
  int32_t* p1 = lds;
  int32_t* p2 = lds + n;
  int32_t* inv1 = lds + 2 * n;
  int32_t* inv2 = lds + 3 * n;
  load_parent(parents + (size_t)r0 * n, n, p1, inv1);
  load_parent(parents + (size_t)(r0 + 1) * n, n, p2, inv2);
  int c1 = cut_lo[q], c2 = cut_hi[q];
  if (c1 > c2) {
    const int t = c1;
    c1 = c2;
    c2 = t;
  }
  c1 = c1 < 0 ? 0 : c1;
  c2 = c2 >= n ? n - 1 : c2;
  __syncthreads();
  int32_t* jump = with_jump ? lds + 4 * n : nullptr;
  pmx_child(p1, p2, inv1, inv2, n, c1, c2, out + (size_t)r0 * n, jump);
  __syncthreads();  // the jump table is reused by the second child
  pmx_child(p2, p1, inv2, inv1, n, c1, c2, out + (size_t)(r0 + 1) * n, jump);
}

// One wave per row.
__global__ __launch_bounds__(kRowBlock) void inversion_kernel(int32_t* __restrict__ routes, int P, int n,
                                                              const double* __restrict__ probs,
                                                              const int32_t* __restrict__ pos1,
                                                              const int32_t* __restrict__ pos2, double rate) {
  const int r = blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= P || !(probs[r] < rate)) return;
  int lo = pos1[r], hi = pos2[r];
  if (lo > hi) {
    const int t = lo;
    lo = hi;
    hi = t;
  }
  lo = lo < 0 ? 0 : lo;
  hi = hi >= n ? n - 1 : hi;
  int32_t* t = routes + (size_t)r * n;
  const int half = (hi - lo + 1) >> 1;
  for (int q = lane; q < half; q += kWave) {
    const int32_t a = t[lo + q];
    const int32_t b = t[hi - q];
    t[lo + q] = b;
    t[hi - q] = a;
  }
}

// Fused tournament -> PMX -> inversion -> evaluate; one workgroup per child row.
__global__ __launch_bounds__(kRowBlock) void generation_kernel(
    const int32_t* __restrict__ pop, const double* __restrict__ fitness, int P, int n, const int32_t* __restrict__ D,
    int ld, const int32_t* __restrict__ tour_cand, int k, const int32_t* __restrict__ cut_lo,
    const int32_t* __restrict__ cut_hi, const double* __restrict__ probs, const int32_t* __restrict__ pos1,
    const int32_t* __restrict__ pos2, double rate, int32_t* __restrict__ offspring, double* __restrict__ off_fit) {
  extern __shared__ int32_t lds[];
  __shared__ int64_t red[kRowBlock / kWave];
  __shared__ int s_w[2];
  generation_child(blockIdx.x, pop, fitness, P, n, MatrixDist{D, ld}, tour_cand, k, cut_lo, cut_hi, probs, pos1,
                   pos2, rate, offspring, off_fit, lds, red, s_w);
}

// Stable ascending order of M float64 keys: order[rank(i)] = i for rank(i) < keep, with
// rank(i) = #{j : (b_j, j) < (b_i, i)} over the orderable key bits b (the total order the bitonic
// network of the batch survival kernel sorts by, so both select the same rows).  A workgroup owns
// KEYS keys; lane L of a wave holds key L mod KEYS and counts over its own share of the j range
// (the 64 / KEYS lane groups of a wave and the 16 waves split it), so a small selection still
// spreads over many CUs: 2048 keys are 256 workgroups of 8 keys, 16 comparisons per lane, where
// 32 workgroups of 64 keys (128 comparisons per lane, a VALU-bound 8.6 us) used 32 CUs.  The keys
// pass through LDS in tiles of kCountTile (already as orderable bits); the lanes of a group read
// the same key j (a broadcast), and the counts are added per key at the end.
constexpr int kCountBlock = 1024;
constexpr int kCountWaves = kCountBlock / kWave;
constexpr int kCountTile = 4096;

// Stable rank of this workgroup's KEYS keys (blockIdx.x * KEYS + lane for lanes < KEYS of wave 0,
// which get the result; other lanes get partial counts).  Every thread must call it.
template <int KEYS>
__device__ __forceinline__ int count_rank(const double* __restrict__ keys, int M, uint64_t* tile,
                                          int (*part)[kWave]) {
  static_assert(kWave % KEYS == 0, "keys per workgroup must divide the wave");
  constexpr int GROUPS = kCountWaves * (kWave / KEYS);  // j-range shares per tile
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int g = wid * (kWave / KEYS) + lane / KEYS;  // this lane's share
  const int i = blockIdx.x * KEYS + (lane & (KEYS - 1));
  const uint64_t bi = i < M ? orderable_bits(keys[i]) : 0ull;
  int cnt = 0;
  for (int base = 0; base < M; base += kCountTile) {
    const int len = (M - base) < kCountTile ? (M - base) : kCountTile;
    __syncthreads();
    for (int x = threadIdx.x; x < len; x += kCountBlock) tile[x] = orderable_bits(keys[base + x]);
    __syncthreads();
    const int chunk = (len + GROUPS - 1) / GROUPS;
    const int j0 = g * chunk;
    const int j1 = (j0 + chunk) < len ? (j0 + chunk) : len;
    const int ib = i - base;  // j < i  <=>  x < ib
#pragma unroll 4
    for (int x = j0; x < j1; ++x) {
      const uint64_t bj = tile[x];
      cnt += (bj < bi) | ((bj == bi) & (x < ib));
    }
  }
#pragma unroll
  for (int m = KEYS; m < kWave; m <<= 1) cnt += __shfl_xor(cnt, m, kWave);  // the wave's groups, per key
  part[wid][lane] = cnt;
  __syncthreads();
  int rank = 0;
#pragma unroll
  for (int w = 0; w < kCountWaves; ++w) rank += part[w][lane & (KEYS - 1)];
  return rank;
}

template <int KEYS>
__global__ __launch_bounds__(kCountBlock) void rank_count_kernel(const double* __restrict__ keys, int M, int keep,
                                                                 int32_t* __restrict__ order) {
  __shared__ uint64_t tile[kCountTile];
  __shared__ int part[kCountWaves][kWave];
  const int rank = count_rank<KEYS>(keys, M, tile, part);
  const int lane = threadIdx.x & (kWave - 1);
  const int i = blockIdx.x * KEYS + lane;
  if (threadIdx.x < KEYS && i < M && rank < keep) order[rank] = i;
}

// (mu + lambda) survival in one launch: the stable rank of every key (as rank_count_kernel), then
// each workgroup copies its own surviving rows into place -- out[rank(i)] = row i of A (i < P0)
// or of B (i >= P0), out_fit[rank(i)] = keys[i], order[rank(i)] = i -- with kCountWaves / KEYS
// waves per row (16-byte copies when `vec`), so the survivors move while the count's workgroups
// are resident instead of in a second launch.
template <int KEYS>
__global__ __launch_bounds__(kCountBlock) void survive_kernel(const double* __restrict__ keys, int M, int keep,
                                                              const int32_t* __restrict__ A, int P0,
                                                              const int32_t* __restrict__ B, int n,
                                                              int32_t* __restrict__ out, double* __restrict__ out_fit,
                                                              int32_t* __restrict__ order, int vec) {
  constexpr int WPR = kCountWaves / KEYS > 0 ? kCountWaves / KEYS : 1;  // waves per row
  __shared__ uint64_t tile[kCountTile];
  __shared__ int part[kCountWaves][kWave];
  __shared__ int s_rank[KEYS];
  const int rank = count_rank<KEYS>(keys, M, tile, part);
  if (threadIdx.x < KEYS) s_rank[threadIdx.x] = blockIdx.x * KEYS + (int)threadIdx.x < M ? rank : keep;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int sub = wid % WPR;
  for (int k = wid / WPR; k < KEYS; k += kCountWaves / WPR) {
    const int r = s_rank[k];
    if (r >= keep) continue;  // uniform per wave
    const int i = blockIdx.x * KEYS + k;
    const int32_t* row = i < P0 ? A + (size_t)i * n : B + (size_t)(i - P0) * n;
    int32_t* dst = out + (size_t)r * n;
    if (vec) {
      const int4* r4 = reinterpret_cast<const int4*>(row);
      int4* o4 = reinterpret_cast<int4*>(dst);
      const int n4 = n >> 2;
      for (int x0 = sub * kWave + lane; x0 < n4; x0 += 4 * WPR * kWave) {
        int4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {  // unconditional (clamped) loads, issued together
          const int x = x0 + u * WPR * kWave;
          v[u] = r4[x < n4 ? x : n4 - 1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (x0 + u * WPR * kWave < n4) o4[x0 + u * WPR * kWave] = v[u];
      }
    } else {
      for (int x = sub * kWave + lane; x < n; x += WPR * kWave) dst[x] = row[x];
    }
    if (sub == 0 && lane == 0) {
      if (out_fit != nullptr) out_fit[r] = keys[i];
      if (order != nullptr) order[r] = i;
    }
  }
}

// One wave per destination row; 16-byte copies when `vec` (n % 4 == 0 and 16-byte aligned rows).
__global__ __launch_bounds__(kRowBlock) void gather_rows_kernel(const int32_t* __restrict__ A,
                                                                const double* __restrict__ fitA, int P0,
                                                                const int32_t* __restrict__ B,
                                                                const double* __restrict__ fitB, int n,
                                                                const int32_t* __restrict__ src, int rows,
                                                                int32_t* __restrict__ out, double* __restrict__ out_fit,
                                                                int vec) {
  const int d = blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (d >= rows) return;
  const int s = src[d];
  const int32_t* row = s < P0 ? A + (size_t)s * n : B + (size_t)(s - P0) * n;
  if (vec) {
    // four 16-byte loads per lane in flight before any store: one memory round trip per 4 KB row
    const int4* r4 = reinterpret_cast<const int4*>(row);
    int4* o4 = reinterpret_cast<int4*>(out + (size_t)d * n);
    const int n4 = n >> 2;
    for (int x0 = lane; x0 < n4; x0 += 4 * kWave) {
      int4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {  // unconditional (clamped) loads, so all four are issued together
        const int x = x0 + u * kWave;
        v[u] = r4[x < n4 ? x : n4 - 1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (x0 + u * kWave < n4) o4[x0 + u * kWave] = v[u];
    }
  } else {
    for (int x = lane; x < n; x += kWave) out[(size_t)d * n + x] = row[x];
  }
  if (out_fit != nullptr && lane == 0) out_fit[d] = s < P0 ? fitA[s] : fitB[s - P0];
}

__global__ __launch_bounds__(kRowBlock) void scatter_rows_kernel(const int32_t* __restrict__ src,
                                                                 const double* __restrict__ src_fit,
                                                                 const int32_t* __restrict__ idx, int rows, int n,
                                                                 int32_t* __restrict__ dst, double* __restrict__ dst_fit) {
  const int r = blockIdx.x * (kRowBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= rows) return;
  const int d = idx[r];
  for (int x = lane; x < n; x += kWave) dst[(size_t)d * n + x] = src[(size_t)r * n + x];
  if (dst_fit != nullptr && src_fit != nullptr && lane == 0) dst_fit[d] = src_fit[r];
}

// Order crossover (OX) of one child per wave: child r of pair q = r / 2 keeps parent r's segment
// [lo, hi] and takes parent r ^ 1's other genes, in that parent's order starting at position
// hi + 1 (wrapping), into positions hi + 1, hi + 2, ... (wrapping) -- order_crossover.py:49-81
// with is_closed false.  The segment's genes are marked in an LDS bitmap; the fill is a wave-wide
// stream compaction (ballot + popcount) over 64 genes at a time.  An odd last row is copied.
constexpr int kOxBlock = 64;

__global__ __launch_bounds__(kOxBlock) void ox_kernel(const int32_t* __restrict__ parents, int P, int n,
                                                      const int32_t* __restrict__ cut_lo,
                                                      const int32_t* __restrict__ cut_hi, int32_t* __restrict__ out) {
  extern __shared__ uint32_t seg_bits[];
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const int32_t* p1 = parents + (size_t)r * n;
  int32_t* child = out + (size_t)r * n;
  if ((P & 1) && r == P - 1) {
    for (int x = lane; x < n; x += kOxBlock) child[x] = p1[x];
    return;
  }
  const int32_t* p2 = parents + (size_t)(r ^ 1) * n;
  const int q = r >> 1;
  int lo = cut_lo[q], hi = cut_hi[q];
  lo = lo < 0 ? 0 : (lo >= n ? n - 1 : lo);
  hi = hi < lo ? lo : (hi >= n ? n - 1 : hi);
  const int words = (n + 31) >> 5;
  for (int w = lane; w < words; w += kOxBlock) seg_bits[w] = 0u;
  __syncthreads();
  for (int x = lo + lane; x <= hi; x += kOxBlock) {
    const int g = clamp_id(p1[x], n);
    child[x] = g;
    atomicOr(&seg_bits[g >> 5], 1u << (g & 31));
  }
  __syncthreads();
  const int shift = hi + 1 == n ? 0 : hi + 1;
  int filled = 0;
  for (int t0 = 0; t0 < n; t0 += kOxBlock) {
    const int t = t0 + lane;
    int g = 0;
    bool keep = false;
    if (t < n) {
      const int src = shift + t < n ? shift + t : shift + t - n;
      g = clamp_id(p2[src], n);
      keep = !((seg_bits[g >> 5] >> (g & 31)) & 1u);
    }
    const uint64_t mask = __ballot(keep);
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    if (keep) {
      int dst = hi + 1 + filled + rank;
      dst = dst >= n ? dst - n : dst;
      child[dst] = g;
    }
    filled += __popcll(mask);
  }
}

// Swap mutation in place: rows with probs[r] < rate exchange positions pos1[r] and pos2[r]
// (swap_mutation.py:84-105).  One thread per row.
__global__ __launch_bounds__(256) void swap_kernel(int32_t* __restrict__ routes, int P, int n,
                                                   const double* __restrict__ probs, const int32_t* __restrict__ pos1,
                                                   const int32_t* __restrict__ pos2, double rate) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P || !(probs[r] < rate)) return;
  const int a = pos1[r], b = pos2[r];
  if ((unsigned)a >= (unsigned)n || (unsigned)b >= (unsigned)n) return;
  int32_t* t = routes + (size_t)r * n;
  const int32_t va = t[a];
  t[a] = t[b];
  t[b] = va;
}

inline int row_blocks(int rows) { return (rows + kRowBlock / kWave - 1) / (kRowBlock / kWave); }

}  // namespace tspmi

using namespace tspmi;

extern "C" int tspmi_tournament(const int32_t* pop, int P, int n, const double* fitness, const int32_t* cand, int k,
                                int num_parents, int32_t* out, int32_t* winners_out, void* stream) {
  TSPMI_CHECK_ARG(P > 0 && n >= 0 && k >= 1 && num_parents >= 0, "tspmi_tournament: bad shape P=%d n=%d k=%d", P, n, k);
  if (num_parents == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(pop && fitness && cand && out, "tspmi_tournament: null pointer");
  hipLaunchKernelGGL(tournament_kernel, dim3(row_blocks(num_parents)), dim3(kRowBlock), 0, (hipStream_t)stream, pop, P,
                     n, fitness, cand, k, num_parents, out, winners_out);
  TSPMI_CHECK_LAUNCH("tspmi_tournament launch");
  return TSPMI_OK;
}

extern "C" int tspmi_pmx(const int32_t* parents, int P, int n, const int32_t* cut_lo, const int32_t* cut_hi,
                         int32_t* out, void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_pmx: negative shape");
  if (P == 0 || n == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(parents && out && (P < 2 || (cut_lo && cut_hi)), "tspmi_pmx: null pointer");
  const size_t lds4 = (size_t)4 * n * sizeof(int32_t);
  TSPMI_CHECK_ARG(lds4 <= (size_t)kMaxLds, "tspmi_pmx: n=%d exceeds the LDS limit", n);
  // The jump table (n more ints) when it fits; above that the chains are walked step by step.
  const int with_jump = lds4 + (size_t)n * sizeof(int32_t) <= (size_t)kMaxLds ? 1 : 0;
  const size_t lds = lds4 + (with_jump ? (size_t)n * sizeof(int32_t) : 0);
  hipLaunchKernelGGL(pmx_kernel, dim3((P + 1) / 2), dim3(kRowBlock), lds, (hipStream_t)stream, parents, P, n, cut_lo,
                     cut_hi, out, with_jump);
  TSPMI_CHECK_LAUNCH("tspmi_pmx launch");
  return TSPMI_OK;
}

extern "C" int tspmi_inversion(int32_t* routes, int P, int n, const double* probs, const int32_t* pos1,
                               const int32_t* pos2, double rate, void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_inversion: negative shape");
  if (P == 0 || n < 2) return TSPMI_OK;
  TSPMI_CHECK_ARG(routes && probs && pos1 && pos2, "tspmi_inversion: null pointer");
  hipLaunchKernelGGL(inversion_kernel, dim3(row_blocks(P)), dim3(kRowBlock), 0, (hipStream_t)stream, routes, P, n,
                     probs, pos1, pos2, rate);
  TSPMI_CHECK_LAUNCH("tspmi_inversion launch");
  return TSPMI_OK;
}

extern "C" int tspmi_generation(const int32_t* pop, const double* fitness, int P, int n, const int32_t* D, int ld,
                                const int32_t* tour_cand, int k, const int32_t* cut_lo, const int32_t* cut_hi,
                                const double* probs, const int32_t* pos1, const int32_t* pos2, double rate,
                                int32_t* offspring, double* off_fit, void* stream) {
  TSPMI_CHECK_ARG(P > 0 && n >= 2 && k >= 1 && ld >= n && (long)n * ld < kMaxMatrixElems,
                  "tspmi_generation: bad shape P=%d n=%d k=%d ld=%d", P, n, k, ld);
  TSPMI_CHECK_ARG(pop && fitness && D && tour_cand && probs && pos1 && pos2 && offspring && off_fit &&
                      (P < 2 || (cut_lo && cut_hi)),
                  "tspmi_generation: null pointer");
  const size_t lds = (size_t)6 * n * sizeof(int32_t);
  TSPMI_CHECK_ARG(lds + 64 <= (size_t)kMaxLds, "tspmi_generation: n=%d exceeds the fused LDS limit", n);
  hipLaunchKernelGGL(generation_kernel, dim3(P), dim3(kRowBlock), lds, (hipStream_t)stream, pop, fitness, P,
                     n, D, ld, tour_cand, k, cut_lo, cut_hi, probs, pos1, pos2, rate, offspring, off_fit);
  TSPMI_CHECK_LAUNCH("tspmi_generation launch");
  return TSPMI_OK;
}

extern "C" int tspmi_rank_select(const double* keys, int M, int keep, int32_t* order, void* stream) {
  TSPMI_CHECK_ARG(M >= 0 && keep >= 0 && keep <= M, "tspmi_rank_select: bad sizes M=%d keep=%d", M, keep);
  if (M == 0 || keep == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(keys && order, "tspmi_rank_select: null pointer");
  TSPMI_CHECK_ARG(M <= (1 << 24), "tspmi_rank_select: M=%d above 2^24", M);
  // small selections: few keys per workgroup, so that the counting spreads over the CUs; large
  // ones: 64 keys per workgroup, so that each workgroup's pass over all M keys serves more of them
  if (M <= 8192)
    hipLaunchKernelGGL(rank_count_kernel<8>, dim3((M + 7) / 8), dim3(kCountBlock), 0, (hipStream_t)stream, keys, M, keep,
                       order);
  else
    hipLaunchKernelGGL(rank_count_kernel<kWave>, dim3((M + kWave - 1) / kWave), dim3(kCountBlock), 0,
                       (hipStream_t)stream, keys, M, keep, order);
  TSPMI_CHECK_LAUNCH("tspmi_rank_select launch");
  return TSPMI_OK;
}

extern "C" int tspmi_survive(const double* keys, int M, int keep, const int32_t* A, int P0, const int32_t* B, int n,
                             int32_t* out, double* out_fit, int32_t* order, void* stream) {
  TSPMI_CHECK_ARG(M >= 0 && keep >= 0 && keep <= M && P0 >= 0 && P0 <= M && n >= 0,
                  "tspmi_survive: bad sizes M=%d keep=%d P0=%d n=%d", M, keep, P0, n);
  if (M == 0 || keep == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(keys && out && (A || P0 == 0) && (B || P0 == M), "tspmi_survive: null pointer");
  TSPMI_CHECK_ARG(M <= (1 << 24), "tspmi_survive: M=%d above 2^24", M);
  const int vec = (n % 4 == 0) && ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && ((uintptr_t)out % 16 == 0);
  if (M <= 8192)
    hipLaunchKernelGGL(survive_kernel<8>, dim3((M + 7) / 8), dim3(kCountBlock), 0, (hipStream_t)stream, keys, M, keep, A,
                       P0, B, n, out, out_fit, order, vec);
  else
    hipLaunchKernelGGL(survive_kernel<kWave>, dim3((M + kWave - 1) / kWave), dim3(kCountBlock), 0, (hipStream_t)stream,
                       keys, M, keep, A, P0, B, n, out, out_fit, order, vec);
  TSPMI_CHECK_LAUNCH("tspmi_survive launch");
  return TSPMI_OK;
}

extern "C" int tspmi_gather_rows(const int32_t* A, const double* fitA, int P0, const int32_t* B, const double* fitB,
                                 int n, const int32_t* src, int rows, int32_t* out, double* out_fit, void* stream) {
  TSPMI_CHECK_ARG(P0 >= 0 && n >= 0 && rows >= 0, "tspmi_gather_rows: negative shape");
  if (rows == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(src && out && (A || P0 == 0), "tspmi_gather_rows: null pointer");
  TSPMI_CHECK_ARG(out_fit == nullptr || (fitA || P0 == 0), "tspmi_gather_rows: out_fit needs source fitness");
  const int vec = (n % 4 == 0) && ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && ((uintptr_t)out % 16 == 0);
  hipLaunchKernelGGL(gather_rows_kernel, dim3(row_blocks(rows)), dim3(kRowBlock), 0, (hipStream_t)stream, A, fitA, P0,
                     B, fitB, n, src, rows, out, out_fit, vec);
  TSPMI_CHECK_LAUNCH("tspmi_gather_rows launch");
  return TSPMI_OK;
}

extern "C" int tspmi_scatter_rows(const int32_t* src, const double* src_fit, const int32_t* idx, int rows, int n,
                                  int32_t* dst, double* dst_fit, void* stream) {
  TSPMI_CHECK_ARG(rows >= 0 && n >= 0, "tspmi_scatter_rows: negative shape");
  if (rows == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(src && idx && dst, "tspmi_scatter_rows: null pointer");
  hipLaunchKernelGGL(scatter_rows_kernel, dim3(row_blocks(rows)), dim3(kRowBlock), 0, (hipStream_t)stream, src, src_fit,
                     idx, rows, n, dst, dst_fit);
  TSPMI_CHECK_LAUNCH("tspmi_scatter_rows launch");
  return TSPMI_OK;
}

extern "C" int tspmi_ox(const int32_t* parents, int P, int n, const int32_t* cut_lo, const int32_t* cut_hi,
                        int32_t* out, void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_ox: negative shape");
  if (P == 0 || n == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(parents && out && (P < 2 || (cut_lo && cut_hi)), "tspmi_ox: null pointer");
  const size_t lds = (size_t)((n + 31) / 32) * sizeof(uint32_t);
  TSPMI_CHECK_ARG(lds <= (size_t)kMaxLds, "tspmi_ox: n=%d exceeds the LDS bitmap", n);
  hipLaunchKernelGGL(ox_kernel, dim3(P), dim3(kOxBlock), lds, (hipStream_t)stream, parents, P, n, cut_lo, cut_hi, out);
  TSPMI_CHECK_LAUNCH("tspmi_ox launch");
  return TSPMI_OK;
}

extern "C" int tspmi_swap(int32_t* routes, int P, int n, const double* probs, const int32_t* pos1, const int32_t* pos2,
                          double rate, void* stream) {
  TSPMI_CHECK_ARG(P >= 0 && n >= 0, "tspmi_swap: negative shape");
  if (P == 0 || n < 2) return TSPMI_OK;
  TSPMI_CHECK_ARG(routes && probs && pos1 && pos2, "tspmi_swap: null pointer");
  hipLaunchKernelGGL(swap_kernel, dim3((P + 255) / 256), dim3(256), 0, (hipStream_t)stream, routes, P, n, probs, pos1,
                     pos2, rate);
  TSPMI_CHECK_LAUNCH("tspmi_swap launch");
  return TSPMI_OK;
}
