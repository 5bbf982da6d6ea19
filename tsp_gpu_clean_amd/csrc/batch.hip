// Multi-instance ("uneven batch") kernels: many TSP instances of different n, each with its own
// population, advanced by one launch per pipeline stage that covers every instance (gfx950).
//
// Instances are described by a device table of int64 records (TSPMI_BATCH_FIELDS per instance):
//   n, P, xy offset (doubles), D offset (int32), packed-table offset (int2), kNN offset (int32),
//   population offset (int32), row offset (sum of earlier P), pair offset (sum of earlier P/2).
// Work is listed on the host as flat (instance, index) arrays, ordered longest-first so the
// persistent grids below balance uneven instances.  Every stage reuses the single-instance device
// bodies, so a batched generation is bit-identical to running each instance alone.
#include <map>
#include <mutex>
#include <utility>

#include "ga_dev.hpp"
#include "two_opt_dev.hpp"

namespace tspmi {

enum : int {
  kBN = 0,
  kBP = 1,
  kBXY = 2,
  kBD = 3,
  kBCT = 4,
  kBCand = 5,
  kBPop = 6,
  kBRow = 7,
  kBPair = 8,
};

struct Inst {
  int n, P;
  int64_t xy, d, ct, cand, pop, row, pair;
};

__device__ __forceinline__ Inst load_inst(const int64_t* __restrict__ desc, int i) {
  const int64_t* r = desc + (size_t)i * TSPMI_BATCH_FIELDS;
  return Inst{(int)r[kBN], (int)r[kBP], r[kBXY], r[kBD], r[kBCT], r[kBCand], r[kBPop], r[kBRow], r[kBPair]};
}

// The instance's coordinates in the form given: int32 pairs (the integer distance form) when
// ixy_all is set, else float64.  Both arrays hold two entries per city at the instance's offset.
__device__ __forceinline__ AnyCoordDist batch_coords(const double* xy_all, const int32_t* ixy_all, const Inst& I) {
  return AnyCoordDist{xy_all != nullptr ? reinterpret_cast<const double2*>(xy_all + I.xy) : nullptr,
                      ixy_all != nullptr ? reinterpret_cast<const int2*>(ixy_all + I.xy) : nullptr};
}

// D rows of every instance: one workgroup per (instance, row u).
__global__ __launch_bounds__(256) void batch_dist_kernel(const int64_t* __restrict__ desc,
                                                         const int32_t* __restrict__ item_inst,
                                                         const int32_t* __restrict__ item_idx, int nitems,
                                                         const double* __restrict__ xy_all, int32_t* __restrict__ D_all) {
  const int it = blockIdx.x;
  if (it >= nitems) return;
  const Inst I = load_inst(desc, item_inst[it]);
  const int u = item_idx[it];
  const double* xy = xy_all + I.xy;
  int32_t* row = D_all + I.d + (size_t)u * I.n;
  const double xu = xy[2 * u], yu = xy[2 * u + 1];
  for (int v = threadIdx.x; v < I.n; v += blockDim.x) {
#pragma clang fp contract(off)
    const double dx = xu - xy[2 * v];
    const double dy = yu - xy[2 * v + 1];
    const double s = dx * dx + dy * dy;
    row[v] = (u == v) ? 0 : (int)(__builtin_sqrt(s) + 0.5);
  }
}

// k-NN rows and packed candidate rows: one workgroup per (instance, row u).
__global__ __launch_bounds__(256) void batch_knn_kernel(const int64_t* __restrict__ desc,
                                                        const int32_t* __restrict__ item_inst,
                                                        const int32_t* __restrict__ item_idx, int nitems, int k,
                                                        const int32_t* __restrict__ D_all, int32_t* __restrict__ cand_all,
                                                        int2* __restrict__ ct_all) {
  __shared__ uint64_t red[256 / kWave];
  const int it = blockIdx.x;
  if (it >= nitems) return;
  const Inst I = load_inst(desc, item_inst[it]);
  const int u = item_idx[it];
  const int n = I.n;
  const int ke = k < n - 1 ? k : n - 1;
  const int32_t* row = D_all + I.d + (size_t)u * n;
  int32_t* out = cand_all + I.cand + (size_t)u * k;
  int2* ct = ct_all + I.ct + (size_t)u * 32;
  uint64_t prev = 0;
  for (int r = 0; r < ke; ++r) {
    uint64_t best = kNoMove;
    for (int v = threadIdx.x; v < n; v += blockDim.x) {
      if (v == u) continue;
      const uint64_t key = ((uint64_t)(uint32_t)row[v] << 32) | (uint32_t)v;
      if ((r == 0 || key > prev) && key < best) best = key;
    }
    best = block_min_u64(best, red);
    if (threadIdx.x == 0) {
      const int nb = (int)(uint32_t)best;
      out[r] = nb;
      ct[r] = make_int2(nb, (int)(best >> 32));
    }
    prev = best;
  }
  for (int r = ke + (int)threadIdx.x; r < 32; r += blockDim.x) {
    if (r < k) out[r] = -1;
    ct[r] = make_int2(-1, 0);
  }
}

// SplitMix64 Fisher-Yates rows: one thread per (instance, row).
__global__ void batch_random_init_kernel(const int64_t* __restrict__ desc, const int32_t* __restrict__ item_inst,
                                         const int32_t* __restrict__ item_idx, int nitems,
                                         const uint64_t* __restrict__ seeds, int32_t* __restrict__ pop) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= nitems) return;
  const Inst I = load_inst(desc, item_inst[it]);
  const int r = item_idx[it];
  int32_t* row = pop + I.pop + (size_t)r * I.n;
  for (int x = 0; x < I.n; ++x) row[x] = x;
  uint64_t s = seeds[I.row + r];
  for (int i = I.n - 1; i > 0; --i) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const int j = (int)(((z >> 32) * (uint64_t)(i + 1)) >> 32);
    const int32_t t = row[i];
    row[i] = row[j];
    row[j] = t;
  }
}

// Tour costs of listed rows into key slots: one wave per (instance, row).
__global__ __launch_bounds__(256) void batch_cost_kernel(const int64_t* __restrict__ desc,
                                                         const int32_t* __restrict__ item_inst,
                                                         const int32_t* __restrict__ item_idx, int nitems,
                                                         const int32_t* __restrict__ pop, const int32_t* __restrict__ D_all,
                                                         const double* __restrict__ xy_all,
                                                         const int32_t* __restrict__ ixy_all, double* __restrict__ keys) {
  const int it = blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (it >= nitems) return;
  const Inst I = load_inst(desc, item_inst[it]);
  const int r = item_idx[it];
  const int32_t* t = pop + I.pop + (size_t)r * I.n;
  const MatrixDist md{D_all + I.d, I.n};
  const AnyCoordDist cd = batch_coords(xy_all, ixy_all, I);
  const bool coords = xy_all != nullptr || ixy_all != nullptr;
  int64_t s = 0;
  for (int x = lane; x < I.n; x += kWave) {
    const int a = clamp_id(t[x], I.n);
    const int b = clamp_id(t[x + 1 == I.n ? 0 : x + 1], I.n);
    if (I.n > 1) s += coords ? cd(a, b) : md(a, b);
  }
  s = wave_sum_i64(s);
  if (lane == 0) keys[2 * I.row + r] = (double)s;
}

// Fused generation of every instance: one workgroup per (instance, child row), persistent grid.
__global__ void batch_generation_kernel(const int64_t* __restrict__ desc, const int32_t* __restrict__ item_inst,
                                        const int32_t* __restrict__ item_idx, int nitems,
                                        const int32_t* __restrict__ popA, double* __restrict__ keysA,
                                        const int32_t* __restrict__ D_all, const double* __restrict__ xy_all,
                                        const int32_t* __restrict__ ixy_all, const int32_t* __restrict__ tour_cand, int k,
                                        const int32_t* __restrict__ cut_lo, const int32_t* __restrict__ cut_hi,
                                        const double* __restrict__ probs, const int32_t* __restrict__ pos1,
                                        const int32_t* __restrict__ pos2, double rate, int32_t* __restrict__ off) {
  extern __shared__ int32_t lds[];
  __shared__ int64_t red[1024 / kWave];
  __shared__ int s_w[2];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const Inst I = load_inst(desc, item_inst[it]);
    double* keys = keysA + 2 * I.row;
    if (xy_all != nullptr || ixy_all != nullptr)
      generation_child(item_idx[it], popA + I.pop, keys, I.P, I.n, batch_coords(xy_all, ixy_all, I),
                       tour_cand + (size_t)I.row * k, k,
                       cut_lo + I.pair, cut_hi + I.pair, probs + I.row, pos1 + I.row, pos2 + I.row, rate, off + I.pop,
                       keys + I.P, lds, red, s_w);
    else
      generation_child(item_idx[it], popA + I.pop, keys, I.P, I.n, MatrixDist{D_all + I.d, I.n},
                       tour_cand + (size_t)I.row * k, k, cut_lo + I.pair, cut_hi + I.pair, probs + I.row, pos1 + I.row,
                       pos2 + I.row, rate, off + I.pop, keys + I.P, lds, red, s_w);
    __syncthreads();
  }
}

// (mu + lambda) survivor order of every instance: one workgroup per instance (2P <= 4096).
__global__ __launch_bounds__(kSortBlock) void batch_survival_kernel(const int64_t* __restrict__ desc, int B,
                                                                    const double* __restrict__ keysA,
                                                                    int32_t* __restrict__ order) {
  __shared__ uint64_t s_key[kSortMax];
  __shared__ int32_t s_idx[kSortMax];
  const int i = blockIdx.x;
  if (i >= B) return;
  const Inst I = load_inst(desc, i);
  const int M = 2 * I.P;
  int M2 = 1;
  while (M2 < M) M2 <<= 1;
  sort_select_body(keysA + 2 * I.row, M, M2, I.P, order + I.row, s_key, s_idx);
}

// Survivor rows into the other buffer, in rank order: one wave per (instance, destination row).
__global__ __launch_bounds__(256) void batch_gather_kernel(const int64_t* __restrict__ desc,
                                                           const int32_t* __restrict__ item_inst,
                                                           const int32_t* __restrict__ item_idx, int nitems,
                                                           const int32_t* __restrict__ popA, const double* __restrict__ keysA,
                                                           const int32_t* __restrict__ off,
                                                           const int32_t* __restrict__ order, int32_t* __restrict__ popB,
                                                           double* __restrict__ keysB) {
  const int it = blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (it >= nitems) return;
  const Inst I = load_inst(desc, item_inst[it]);
  const int d = item_idx[it];
  const int s = order[I.row + d];
  const int32_t* src = (s < I.P ? popA + I.pop + (size_t)s * I.n : off + I.pop + (size_t)(s - I.P) * I.n);
  int32_t* dst = popB + I.pop + (size_t)d * I.n;
  for (int x = lane; x < I.n; x += kWave) dst[x] = src[x];
  if (lane == 0) keysB[2 * I.row + d] = keysA[2 * I.row + s];
}

// Elite candidate 2-opt of every instance: persistent grid over (instance, elite row) items.
template <int BLOCK, int KL>
__global__ __launch_bounds__(BLOCK) void batch_two_opt_kernel(const int64_t* __restrict__ desc,
                                                              const int32_t* __restrict__ item_inst,
                                                              const int32_t* __restrict__ item_idx, int nitems,
                                                              int32_t* __restrict__ pop, const int32_t* __restrict__ D_all,
                                                              const int2* __restrict__ ct_all, int k, int max_iter,
                                                              int64_t* __restrict__ stats) {
  extern __shared__ int32_t lds[];
  __shared__ uint64_t red[BLOCK / kWave];
  __shared__ int64_t red64[BLOCK / kWave];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const Inst I = load_inst(desc, item_inst[it]);
    const int ke = k < I.n - 1 ? k : I.n - 1;
    if (I.n >= 4 && ke >= 1)
      cand_two_opt_tour<BLOCK, 32, 4, KL>(pop + I.pop + (size_t)item_idx[it] * I.n, I.n,
                                          MatrixDist{D_all + I.d, I.n}, ct_all + I.ct, ke, max_iter, stats, lds, red,
                                          red64);
    __syncthreads();
  }
}

// As batch_two_opt_kernel, but instances with n >= coord_min_n compute their distances from the
// EUC_2D coordinates (same tours) instead of gathering them from a matrix that no longer fits the
// caches; with xy_lds, int32-coordinate instances take rec_two_opt_tour (records in LDS).
template <int BLOCK, int KL>
__global__ __launch_bounds__(BLOCK) void batch_two_opt_mixed_kernel(
    const int64_t* __restrict__ desc, const int32_t* __restrict__ item_inst, const int32_t* __restrict__ item_idx,
    int nitems, int32_t* __restrict__ pop, const int32_t* __restrict__ D_all, const double* __restrict__ xy_all,
    const int32_t* __restrict__ ixy_all, const int2* __restrict__ ct_all, int k, int max_iter, int coord_min_n,
    int xy_lds, int64_t* __restrict__ stats) {
  extern __shared__ __align__(16) int32_t lds[];
  __shared__ uint64_t red[BLOCK / kWave];
  __shared__ int64_t red64[BLOCK / kWave];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const Inst I = load_inst(desc, item_inst[it]);
    const int ke = k < I.n - 1 ? k : I.n - 1;
    int32_t* tour = pop + I.pop + (size_t)item_idx[it] * I.n;
    if (I.n >= 4 && ke >= 1) {
      if (I.n >= coord_min_n && ixy_all != nullptr) {
        const int2* ixy = reinterpret_cast<const int2*>(ixy_all + I.xy);
        if (xy_lds)
          rec_two_opt_tour<BLOCK, 32, 4, KL>(tour, I.n, ixy, ct_all + I.ct, ke, max_iter, stats, lds, red, red64);
        else
          cand_two_opt_tour<BLOCK, 32, 4, KL>(tour, I.n, IntCoordDist{ixy}, ct_all + I.ct, ke, max_iter, stats, lds,
                                              red, red64);
      }
      else if (I.n >= coord_min_n)
        cand_two_opt_tour<BLOCK, 32, 4, KL>(tour, I.n, CoordDist{reinterpret_cast<const double2*>(xy_all + I.xy)},
                                            ct_all + I.ct, ke, max_iter, stats, lds, red, red64);
      else
        cand_two_opt_tour<BLOCK, 32, 4, KL>(tour, I.n, MatrixDist{D_all + I.d, I.n}, ct_all + I.ct, ke, max_iter,
                                            stats, lds, red, red64);
    }
    __syncthreads();
  }
}

// Index of the first minimum of keys[0 .. c) for a whole wave (returned to every lane).
__device__ __forceinline__ int wave_first_min(const double* __restrict__ keys, int c, int lane) {
  double mv = 0.0;
  int mi = 0x7fffffff;
  for (int x = lane; x < c; x += kWave) {
    const double v = keys[x];
    if (mi == 0x7fffffff || v < mv) {
      mv = v;
      mi = x;
    }
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double ov = __shfl_xor(mv, m, kWave);
    const int oi = __shfl_xor(mi, m, kWave);
    if (oi != 0x7fffffff && (mi == 0x7fffffff || ov < mv || (ov == mv && oi < mi))) {
      mv = ov;
      mi = oi;
    }
  }
  return mi;
}

// First minimum of the first `count` keys of every instance: one wave per instance.
__global__ __launch_bounds__(256) void batch_best_kernel(const int64_t* __restrict__ desc, int B,
                                                         const double* __restrict__ keys, int count,
                                                         int32_t* __restrict__ best) {
  const int i = blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (i >= B) return;
  const Inst I = load_inst(desc, i);
  const int c = count < I.P ? count : I.P;
  const int mi = wave_first_min(keys + 2 * I.row, c, lane);
  if (lane == 0) best[i] = mi;
}

// ---- one persistent launch per generation ---------------------------------------------------
//
// A generation of every instance is a list of tasks in an order where each task's inputs are
// produced by tasks earlier in the list:
//   child (i, g)   : tournament -> PMX -> inversion -> cost of child rows g*cpt .. g*cpt + cpt - 1
//                    of instance i (cpt even, so a crossover pair stays in one task);
//   survive (i)    : after every child task of i: the stable (mu + lambda) rank of its 2P keys;
//                    publishes the elite rows' sources, then copies the other survivors into the
//                    next population with their keys;
//   elite (i, d)   : after survive (i): survivor d (d < elite) copied into the next population and
//                    refined by candidate 2-opt, with its cost.
// The host lists all child tasks (instances longest first), then every survive task, then the
// elite tasks (longest first).  Workgroups draw tasks in list order from one ticket counter, so
// every task a workgroup waits for was drawn earlier by a workgroup that is running: the smallest
// unfinished task never waits, and the launch always drains whatever the number of resident
// workgroups.  A task publishes its results with one agent-scope release: every wave first waits
// for its own stores to complete, then after the workgroup barrier one thread writes the L2 back
// and updates the counter (a release per wave would write the L2 back once per wave, which
// measured several times slower).  A waiting task polls with relaxed loads and takes an
// agent-scope acquire before the workgroup reads what its producers wrote.  Survivor readiness is
// an epoch (launches counted on the device, so graph replays never see a stale flag); child
// counts are cleared by the survive task that consumed them, the elite completion count (kept
// only when the best row is asked for) by the last elite task, and the ticket counter by the last
// workgroup to leave.  Waits are bounded: one that runs out sets the error word, after which no
// task of the launch waits, and the caller must treat the call as failed.
enum : uint32_t { kTaskChild = 0, kTaskSurvive = 1, kTaskElite = 2 };
constexpr int kFusedHeader = 4;  // ticket, exited, error, epoch; then 3 words per instance
constexpr int kFusedSpinLimit = 1 << 20;  // polls of ~150 clocks: about 65 ms
constexpr uint32_t kNoTask = 0xffffffffu;

__device__ __forceinline__ uint32_t task_kind(uint32_t t) { return t >> 30; }
__device__ __forceinline__ int task_inst(uint32_t t) { return (int)((t >> 12) & 0x3ffffu); }
__device__ __forceinline__ int task_index(uint32_t t) { return (int)(t & 0xfffu); }

// Thread 0 waits until *word >= target, then acquires at agent scope; the barrier hands that
// ordering to the whole workgroup.  Once any wait of the launch has run out, no task waits any
// more, so a failed launch still ends within about one wait limit.
__device__ __forceinline__ void task_wait(int* word, int target, int* err) {
  if (threadIdx.x == 0) {
    int spin = 0;
    while (__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spin > kFusedSpinLimit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Every wave's global stores have completed (reached the L2) when the workgroup leaves this
// barrier; a workgroup-scope barrier alone does not wait for them.
__device__ __forceinline__ void stores_complete_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
}

// After stores_complete_barrier: thread 0 adds 1 to *word with an agent-scope release (or, with
// ACQUIRE, an acquire-release that also makes every earlier publisher's stores visible here) and
// gets the old value (other threads get 0).
template <bool ACQUIRE = false>
__device__ __forceinline__ int task_publish_add(int* word) {
  stores_complete_barrier();
  int old = 0;
  if (threadIdx.x == 0)
    old = __hip_atomic_fetch_add(word, 1, ACQUIRE ? __ATOMIC_ACQ_REL : __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  return old;
}

// Copies survivor rows [d0, d1) of an instance (sources from `ord`, in LDS) with their keys.
__device__ __forceinline__ void copy_survivors(const Inst& I, const int32_t* ord, int d0, int d1,
                                               const int32_t* __restrict__ popA, const double* __restrict__ keysA,
                                               const int32_t* __restrict__ off, int32_t* __restrict__ popB,
                                               double* __restrict__ keysB) {
  const int n = I.n;
  const int nw = blockDim.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  constexpr int U = 8;  // loads in flight per lane before any store
  // one wave per row, so a row's loads and stores stay in one wave's issue stream
  for (int d = d0 + (int)threadIdx.x / kWave; d < d1; d += nw) {
    const int s = ord[d];
    const int32_t* src = s < I.P ? popA + I.pop + (size_t)s * n : off + I.pop + (size_t)(s - I.P) * n;
    int32_t* dst = popB + I.pop + (size_t)d * n;
    for (int x0 = lane; x0 < n; x0 += U * kWave) {
      int v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = x0 + u * kWave;
        v[u] = x < n ? src[x] : 0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = x0 + u * kWave;
        if (x < n) dst[x] = v[u];
      }
    }
    if (lane == 0) keysB[2 * I.row + d] = keysA[2 * I.row + s];
  }
}

template <int BLOCK, int KL>
__global__ __launch_bounds__(BLOCK) void batch_fused_generation_kernel(
    const int64_t* __restrict__ desc, int B, const uint32_t* __restrict__ tasks, int ntasks, int* __restrict__ ctr,
    int cpt, const int32_t* __restrict__ popA, double* __restrict__ keysA, int32_t* __restrict__ off,
    int32_t* __restrict__ order, int32_t* __restrict__ popB, double* __restrict__ keysB,
    const int32_t* __restrict__ D_all, const double* __restrict__ xy_all, const int32_t* __restrict__ ixy_all,
    const int2* __restrict__ ct_all, const int32_t* __restrict__ tour_cand, int k_tour, const int32_t* __restrict__ cut_lo,
    const int32_t* __restrict__ cut_hi, const double* __restrict__ probs, const int32_t* __restrict__ pos1,
    const int32_t* __restrict__ pos2, double rate, int elite, int kc, int max_iter, int coord_min_n,
    int64_t* __restrict__ stats, int32_t* __restrict__ best, int best_count) {
  extern __shared__ __align__(16) int32_t lds[];
  __shared__ uint64_t red[BLOCK / kWave];
  __shared__ int64_t red64[BLOCK / kWave];
  __shared__ int s_w[2];
  __shared__ uint32_t s_task;
  __shared__ int s_epoch, s_flag;
  int* const head = ctr;
  int* const exited = ctr + 1;
  int* const err = ctr + 2;
  int* const child_done = ctr + kFusedHeader;
  int* const ready = child_done + B;
  int* const elite_done = ready + B;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  if (tid == 0) s_epoch = __hip_atomic_load(ctr + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
  for (;;) {
    if (tid == 0) {
      const int t = atomicAdd(head, 1);
      s_task = t < ntasks ? tasks[t] : kNoTask;
    }
    __syncthreads();
    const uint32_t task = s_task;
    if (task == kNoTask) break;
    const int inst = task_inst(task);
    const int idx = task_index(task);
    const Inst I = load_inst(desc, inst);
    const int n = I.n;
    const int e = elite < I.P ? elite : I.P;
    const AnyCoordDist cd = batch_coords(xy_all, ixy_all, I);
    const bool coords = xy_all != nullptr || ixy_all != nullptr;
    const MatrixDist md{D_all + I.d, n};
    const uint32_t kind = task_kind(task);
    if (kind == kTaskChild) {
      double* keys = keysA + 2 * I.row;
      const int r1 = (idx + 1) * cpt < I.P ? (idx + 1) * cpt : I.P;
      for (int r = idx * cpt; r < r1; ++r) {
        if (coords)
          generation_child(r, popA + I.pop, keys, I.P, n, cd, tour_cand + (size_t)I.row * k_tour, k_tour,
                           cut_lo + I.pair, cut_hi + I.pair, probs + I.row, pos1 + I.row, pos2 + I.row, rate,
                           off + I.pop, keys + I.P, lds, red64, s_w);
        else
          generation_child(r, popA + I.pop, keys, I.P, n, md, tour_cand + (size_t)I.row * k_tour, k_tour,
                           cut_lo + I.pair, cut_hi + I.pair, probs + I.row, pos1 + I.row, pos2 + I.row, rate,
                           off + I.pop, keys + I.P, lds, red64, s_w);
        __syncthreads();  // the next child reuses the LDS
      }
      task_publish_add(child_done + inst);
    } else if (kind == kTaskSurvive) {
      task_wait(child_done + inst, (I.P + cpt - 1) / cpt, err);
      // rank(x) = #{j : (b_j, j) < (b_x, x)} over orderable key bits, as rank_count_kernel
      const int M = 2 * I.P;
      uint32_t* s_hi = reinterpret_cast<uint32_t*>(lds);
      uint32_t* s_lo = s_hi + M;
      int32_t* s_ord = reinterpret_cast<int32_t*>(s_lo + M);
      const double* keys = keysA + 2 * I.row;
      for (int x = tid; x < M; x += BLOCK) {
        const uint64_t b = orderable_bits(keys[x]);
        s_hi[x] = (uint32_t)(b >> 32);
        s_lo[x] = (uint32_t)b;
      }
      __syncthreads();
      for (int x = tid; x < M; x += BLOCK) {
        const uint64_t bx = ((uint64_t)s_hi[x] << 32) | s_lo[x];
        int rank = 0;
        for (int j = 0; j < M; ++j) {
          const uint64_t bj = ((uint64_t)s_hi[j] << 32) | s_lo[j];
          rank += (bj < bx) | ((bj == bx) & (j < x));
        }
        if (rank < I.P) {
          s_ord[rank] = x;
          if (rank < e) order[I.row + rank] = x;
        }
      }
      // the elite tasks need only their sources: release them before copying the other survivors
      if (tid == 0) __hip_atomic_store(child_done + inst, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      stores_complete_barrier();
      if (tid == 0) __hip_atomic_store(ready + inst, s_epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      copy_survivors(I, s_ord, e, I.P, popA, keysA, off, popB, keysB);
      __syncthreads();
      if (e == 0 && best != nullptr && tid < kWave) {  // no elite task follows: the best row is here
        const int c = best_count < I.P ? best_count : I.P;
        const int mi = wave_first_min(keysB + 2 * I.row, c, lane);
        if (lane == 0) best[inst] = mi;
      }
    } else {  // elite row idx
      task_wait(ready + inst, s_epoch, err);
      const int s = order[I.row + idx];
      const int32_t* src = s < I.P ? popA + I.pop + (size_t)s * n : off + I.pop + (size_t)(s - I.P) * n;
      int32_t* dst = popB + I.pop + (size_t)idx * n;
      double* kout = keysB + 2 * I.row + idx;
      const int ke = kc < n - 1 ? kc : n - 1;
      if (n >= 4 && ke >= 1 && max_iter != 0) {
        if (n >= coord_min_n && ixy_all != nullptr)
          // coordinates staged in LDS after the scan's arrays (5n + 2 words, within the child task's 6n);
          // rec_two_opt_tour measured slower inside this kernel (300 against 670 generations/s)
          cand_two_opt_tour<BLOCK, 32, 4, KL>(
              dst, n, stage_int_coords<BLOCK>(reinterpret_cast<const int2*>(ixy_all + I.xy), n, lds), ct_all + I.ct,
              ke, max_iter, stats, lds, red, red64, kout, nullptr, 0, src);
        else if (n >= coord_min_n)
          cand_two_opt_tour<BLOCK, 32, 4, KL>(dst, n, CoordDist{reinterpret_cast<const double2*>(xy_all + I.xy)},
                                              ct_all + I.ct, ke, max_iter, stats, lds, red, red64, kout, nullptr, 0,
                                              src);
        else
          cand_two_opt_tour<BLOCK, 32, 4, KL>(dst, n, md, ct_all + I.ct, ke, max_iter, stats, lds, red, red64, kout,
                                              nullptr, 0, src);
      } else {  // no refinement: copy, and cost the row as tspmi_batch_cost does
        int64_t part = 0;
        for (int x = tid; x < n; x += BLOCK) {
          dst[x] = src[x];
          const int a = clamp_id(src[x], n);
          const int b = clamp_id(src[x + 1 == n ? 0 : x + 1], n);
          if (n > 1) part += coords ? cd(a, b) : md(a, b);
        }
        const int64_t total = block_sum_i64(part, red64);
        if (tid == 0) *kout = (double)total;
      }
      if (best != nullptr) {  // the instance's last elite task takes the first minimum of its keys
        const int old = task_publish_add<true>(elite_done + inst);
        if (tid == 0) s_flag = (old == e - 1);
        __syncthreads();
        if (s_flag) {
          if (tid == 0) __hip_atomic_store(elite_done + inst, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (tid < kWave) {
            const int c = best_count < I.P ? best_count : I.P;
            const int mi = wave_first_min(keysB + 2 * I.row, c, lane);
            if (lane == 0) best[inst] = mi;
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    // the last workgroup to leave advances the epoch and clears the ticket counter
    if (__hip_atomic_fetch_add(exited, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == (int)gridDim.x - 1) {
      __hip_atomic_store(ctr + 3, s_epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(head, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(exited, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace tspmi

using namespace tspmi;

#define TSPMI_BATCH_ARGS(nm)                                                      \
  TSPMI_CHECK_ARG(desc != nullptr && nitems >= 0 && (nitems == 0 || (item_inst && item_idx)), \
                  nm ": bad work list")

extern "C" int tspmi_batch_dist_euc2d(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx,
                                      int nitems, const double* xy_all, int32_t* D_all, void* stream) {
  TSPMI_BATCH_ARGS("tspmi_batch_dist_euc2d");
  if (nitems == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(xy_all && D_all, "tspmi_batch_dist_euc2d: null pointer");
  hipLaunchKernelGGL(batch_dist_kernel, dim3(nitems), dim3(256), 0, (hipStream_t)stream, desc, item_inst, item_idx,
                     nitems, xy_all, D_all);
  TSPMI_CHECK_LAUNCH("tspmi_batch_dist_euc2d launch");
  return TSPMI_OK;
}

extern "C" int tspmi_batch_knn(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems,
                               int k, const int32_t* D_all, int32_t* cand_all, int32_t* table_all, void* stream) {
  TSPMI_BATCH_ARGS("tspmi_batch_knn");
  TSPMI_CHECK_ARG(k >= 1 && k <= 32, "tspmi_batch_knn: k=%d outside [1, 32]", k);
  if (nitems == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(D_all && cand_all && table_all, "tspmi_batch_knn: null pointer");
  hipLaunchKernelGGL(batch_knn_kernel, dim3(nitems), dim3(256), 0, (hipStream_t)stream, desc, item_inst, item_idx,
                     nitems, k, D_all, cand_all, reinterpret_cast<int2*>(table_all));
  TSPMI_CHECK_LAUNCH("tspmi_batch_knn launch");
  return TSPMI_OK;
}

extern "C" int tspmi_batch_random_init(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx,
                                       int nitems, const uint64_t* seeds, int32_t* pop, void* stream) {
  TSPMI_BATCH_ARGS("tspmi_batch_random_init");
  if (nitems == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(seeds && pop, "tspmi_batch_random_init: null pointer");
  hipLaunchKernelGGL(batch_random_init_kernel, dim3((nitems + 63) / 64), dim3(64), 0, (hipStream_t)stream, desc,
                     item_inst, item_idx, nitems, seeds, pop);
  TSPMI_CHECK_LAUNCH("tspmi_batch_random_init launch");
  return TSPMI_OK;
}

extern "C" int tspmi_batch_cost(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems,
                                const int32_t* pop, const int32_t* D_all, const double* xy_all, const int32_t* ixy_all,
                                double* keys, void* stream) {
  TSPMI_BATCH_ARGS("tspmi_batch_cost");
  if (nitems == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(pop && (D_all || xy_all || ixy_all) && keys, "tspmi_batch_cost: null pointer");
  TSPMI_CHECK_ARG(((uintptr_t)xy_all & 15) == 0 && ((uintptr_t)ixy_all & 7) == 0,
                  "tspmi_batch_cost: xy must be 16-byte and ixy 8-byte aligned");
  hipLaunchKernelGGL(batch_cost_kernel, dim3((nitems + 3) / 4), dim3(256), 0, (hipStream_t)stream, desc, item_inst,
                     item_idx, nitems, pop, D_all, xy_all, ixy_all, keys);
  TSPMI_CHECK_LAUNCH("tspmi_batch_cost launch");
  return TSPMI_OK;
}

extern "C" int tspmi_batch_generation(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx,
                                      int nitems, int max_n, const int32_t* popA, double* keysA, const int32_t* D_all,
                                      const double* xy_all, const int32_t* ixy_all, const int32_t* tour_cand, int k,
                                      const int32_t* cut_lo,
                                      const int32_t* cut_hi,
                                      const double* probs, const int32_t* pos1, const int32_t* pos2, double rate,
                                      int32_t* off, int block, int grid, void* stream) {
  TSPMI_BATCH_ARGS("tspmi_batch_generation");
  TSPMI_CHECK_ARG(block == 128 || block == 256 || block == 512 || block == 1024,
                  "tspmi_batch_generation: block %d not in {128, 256, 512, 1024}", block);
  if (nitems == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(((uintptr_t)xy_all & 15) == 0 && ((uintptr_t)ixy_all & 7) == 0,
                  "tspmi_batch_generation: xy must be 16-byte and ixy 8-byte aligned");
  TSPMI_CHECK_ARG(popA && keysA && (D_all || xy_all || ixy_all) && tour_cand && cut_lo && cut_hi && probs && pos1 && pos2 && off &&
                      k >= 1,
                  "tspmi_batch_generation: null pointer");
  const size_t lds = (size_t)6 * max_n * sizeof(int32_t);
  TSPMI_CHECK_ARG(max_n >= 2 && lds + 1024 <= (size_t)kMaxLds, "tspmi_batch_generation: max_n=%d exceeds LDS", max_n);
  const int g = grid > 0 && grid < nitems ? grid : nitems;
  hipLaunchKernelGGL(batch_generation_kernel, dim3(g), dim3(block), lds, (hipStream_t)stream, desc, item_inst, item_idx,
                     nitems, popA, keysA, D_all, xy_all, ixy_all, tour_cand, k, cut_lo, cut_hi, probs, pos1, pos2, rate,
                     off);
  TSPMI_CHECK_LAUNCH("tspmi_batch_generation launch");
  return TSPMI_OK;
}

extern "C" int tspmi_batch_survival(const int64_t* desc, int B, const double* keysA, int32_t* order, void* stream) {
  TSPMI_CHECK_ARG(desc && B >= 0, "tspmi_batch_survival: bad arguments");
  if (B == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(keysA && order, "tspmi_batch_survival: null pointer");
  hipLaunchKernelGGL(batch_survival_kernel, dim3(B), dim3(kSortBlock), 0, (hipStream_t)stream, desc, B, keysA, order);
  TSPMI_CHECK_LAUNCH("tspmi_batch_survival launch");
  return TSPMI_OK;
}

extern "C" int tspmi_batch_gather(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems,
                                  const int32_t* popA, const double* keysA, const int32_t* off, const int32_t* order,
                                  int32_t* popB, double* keysB, void* stream) {
  TSPMI_BATCH_ARGS("tspmi_batch_gather");
  if (nitems == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(popA && keysA && off && order && popB && keysB, "tspmi_batch_gather: null pointer");
  hipLaunchKernelGGL(batch_gather_kernel, dim3((nitems + 3) / 4), dim3(256), 0, (hipStream_t)stream, desc, item_inst,
                     item_idx, nitems, popA, keysA, off, order, popB, keysB);
  TSPMI_CHECK_LAUNCH("tspmi_batch_gather launch");
  return TSPMI_OK;
}

extern "C" int tspmi_batch_two_opt_packed(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx,
                                          int nitems, int max_n, int32_t* pop, const int32_t* D_all,
                                          const int32_t* table_all, int k, int max_iter, int block, int grid,
                                          int64_t* stats, void* stream) {
  TSPMI_BATCH_ARGS("tspmi_batch_two_opt_packed");
  TSPMI_CHECK_ARG(k >= 1 && k <= 32, "tspmi_batch_two_opt_packed: k=%d outside [1, 32]", k);
  if (nitems == 0 || max_iter == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(pop && D_all && table_all, "tspmi_batch_two_opt_packed: null pointer");
  const size_t lds = (size_t)(3 * max_n + 1) * sizeof(int32_t);
  TSPMI_CHECK_ARG(lds + 2048 <= (size_t)kMaxLds, "tspmi_batch_two_opt_packed: max_n=%d exceeds LDS", max_n);
  const int g = grid > 0 && grid < nitems ? grid : nitems;
  hipStream_t s = (hipStream_t)stream;
  const int2* ct = reinterpret_cast<const int2*>(table_all);
#define TSPMI_BATCH_LS(BL, KPV, KLV)                                                                          \
  hipLaunchKernelGGL((batch_two_opt_kernel<BL, KLV>), dim3(g), dim3(BL), lds, s, desc, item_inst, item_idx, nitems, \
                     pop, D_all, ct, k, max_iter, stats)
  if (block == 256) {
    TSPMI_CAND_BY_K(TSPMI_BATCH_LS, 256, k)
  } else if (block == 512) {
    TSPMI_CAND_BY_K(TSPMI_BATCH_LS, 512, k)
  } else if (block == 1024) {
    TSPMI_CAND_BY_K(TSPMI_BATCH_LS, 1024, k)
  } else
    TSPMI_CHECK_ARG(false, "tspmi_batch_two_opt_packed: block %d not in {256, 512, 1024}", block);
  TSPMI_CHECK_LAUNCH("tspmi_batch_two_opt_packed launch");
  return TSPMI_OK;
}

extern "C" int tspmi_batch_best(const int64_t* desc, int B, const double* keys, int count, int32_t* best,
                                void* stream) {
  TSPMI_CHECK_ARG(desc && B >= 0 && count >= 1 && count <= 4096, "tspmi_batch_best: bad arguments");
  if (B == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(keys && best, "tspmi_batch_best: null pointer");
  hipLaunchKernelGGL(batch_best_kernel, dim3((B + 3) / 4), dim3(256), 0, (hipStream_t)stream, desc, B, keys, count,
                     best);
  TSPMI_CHECK_LAUNCH("tspmi_batch_best launch");
  return TSPMI_OK;
}

extern "C" int tspmi_batch_two_opt_packed_mixed(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx,
                                                int nitems, int max_n, int32_t* pop, const int32_t* D_all,
                                                const double* xy_all, const int32_t* ixy_all, const int32_t* table_all,
                                                int k, int max_iter, int block, int grid, int coord_min_n,
                                                int64_t* stats, void* stream) {
  TSPMI_CHECK_ARG(desc && nitems >= 0 && max_n >= 0 && k >= 1 && k <= 32, "tspmi_batch_two_opt_packed_mixed: bad args");
  if (nitems == 0 || max_iter == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(item_inst && item_idx && pop && D_all && (xy_all || ixy_all) && table_all,
                  "tspmi_batch_two_opt_packed_mixed: null pointer");
  TSPMI_CHECK_ARG(((uintptr_t)xy_all & 15) == 0 && ((uintptr_t)ixy_all & 7) == 0,
                  "tspmi_batch_two_opt_packed_mixed: xy must be 16-byte and ixy 8-byte aligned");
  size_t lds = (size_t)(3 * max_n + 1) * sizeof(int32_t);
  TSPMI_CHECK_ARG(lds + 1024 <= (size_t)kMaxLds, "tspmi_batch_two_opt_packed_mixed: n=%d exceeds the LDS limit",
                  max_n);
  const size_t lds_xy = (size_t)rec_lds_words(max_n) * sizeof(int32_t);
  const int xy_lds = ixy_all != nullptr && lds_xy + 1024 <= (size_t)kMaxLds;
  if (xy_lds) lds = lds_xy;
  const int g = grid > 0 && grid < nitems ? grid : nitems;
  hipStream_t s = (hipStream_t)stream;
  const int2* ct = reinterpret_cast<const int2*>(table_all);
#define TSPMI_BATCH_MIXED(BL, KPV, KLV)                                                                        \
  hipLaunchKernelGGL((batch_two_opt_mixed_kernel<BL, KLV>), dim3(g), dim3(BL), lds, s, desc, item_inst, item_idx,   \
                     nitems, pop, D_all, xy_all, ixy_all, ct, k, max_iter, coord_min_n, xy_lds, stats)
  if (block == 256) {
    TSPMI_CAND_BY_K(TSPMI_BATCH_MIXED, 256, k)
  } else if (block == 512) {
    TSPMI_CAND_BY_K(TSPMI_BATCH_MIXED, 512, k)
  } else if (block == 1024) {
    TSPMI_CAND_BY_K(TSPMI_BATCH_MIXED, 1024, k)
  } else
    TSPMI_CHECK_ARG(false, "tspmi_batch_two_opt_packed_mixed: block %d not in {256, 512, 1024}", block);
  TSPMI_CHECK_LAUNCH("tspmi_batch_two_opt_packed_mixed launch");
  return TSPMI_OK;
}

namespace {

template <int BL, int KLV>
int fused_launch(int grid, size_t lds, hipStream_t s, const int64_t* desc, int B, const uint32_t* tasks, int ntasks,
                 int32_t* counters, int cpt, const int32_t* popA, double* keysA, int32_t* off, int32_t* order, int32_t* popB,
                 double* keysB, const int32_t* D_all, const double* xy_all, const int32_t* ixy_all, const int2* ct,
                 const int32_t* tour_cand,
                 int k_tour, const int32_t* cut_lo, const int32_t* cut_hi, const double* probs, const int32_t* pos1,
                 const int32_t* pos2, double rate, int elite, int k, int max_iter, int coord_min_n, int64_t* stats,
                 int32_t* best, int best_count) {
  auto kern = batch_fused_generation_kernel<BL, KLV>;
  if (grid <= 0) {
    // as many workgroups as can be resident at once (more would only find the list drained),
    // queried once per device and LDS size so that a graph capture makes no runtime query
    static std::mutex mu;
    static std::map<std::pair<int, size_t>, int> cache;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "tspmi_batch_fused_generation device");
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({dev, lds});
    if (it == cache.end()) {
      int per_cu = 0;
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BL, lds);
      if (e != hipSuccess) return hip_fail(e, "tspmi_batch_fused_generation occupancy");
      it = cache.emplace(std::make_pair(dev, lds), (per_cu > 0 ? per_cu : 1) * device_cu_count()).first;
    }
    grid = it->second;
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(BL), lds, s, desc, B, tasks, ntasks, counters, cpt, popA, keysA, off, order,
                     popB, keysB, D_all, xy_all, ixy_all, ct, tour_cand, k_tour, cut_lo, cut_hi, probs, pos1, pos2, rate,
                     elite,
                     k, max_iter, coord_min_n, stats, best, best_count);
  TSPMI_CHECK_LAUNCH("tspmi_batch_fused_generation launch");
  return TSPMI_OK;
}

}  // namespace

extern "C" int tspmi_batch_fused_generation(
    const int64_t* desc, int B, const uint32_t* tasks, int ntasks, int32_t* counters, int children_per_task,
    int max_n, int max_p, const int32_t* popA, double* keysA, int32_t* off, int32_t* order, int32_t* popB, double* keysB,
    const int32_t* D_all, const double* xy_all, const int32_t* ixy_all, const int32_t* table_all,
    const int32_t* tour_cand, int k_tour,
    const int32_t* cut_lo, const int32_t* cut_hi, const double* probs, const int32_t* pos1, const int32_t* pos2,
    double rate, int elite, int k, int max_iter, int coord_min_n, int block, int grid, int64_t* stats, int32_t* best,
    int best_count, void* stream) {
  TSPMI_CHECK_ARG(desc && B >= 0 && ntasks >= 0 && (ntasks == 0 || (tasks && counters)),
                  "tspmi_batch_fused_generation: bad task list");
  if (ntasks == 0) return TSPMI_OK;
  TSPMI_CHECK_ARG(children_per_task >= 2 && children_per_task % 2 == 0,
                  "tspmi_batch_fused_generation: children_per_task=%d must be even and >= 2", children_per_task);
  TSPMI_CHECK_ARG(B < (1 << 18) && max_p >= 1 && 2 * max_p <= 4096 && k_tour >= 1 && k >= 1 && k <= 32 &&
                      elite >= 0 && best_count >= 1,
                  "tspmi_batch_fused_generation: B=%d P=%d k_tour=%d k=%d elite=%d outside the supported range", B,
                  max_p, k_tour, k, elite);
  TSPMI_CHECK_ARG(popA && keysA && off && order && popB && keysB && (D_all || xy_all || ixy_all) && table_all &&
                      tour_cand &&
                      cut_lo && cut_hi && probs && pos1 && pos2 && stats,
                  "tspmi_batch_fused_generation: null pointer");
  TSPMI_CHECK_ARG(((uintptr_t)xy_all & 15) == 0 && ((uintptr_t)ixy_all & 7) == 0,
                  "tspmi_batch_fused_generation: xy must be 16-byte and ixy 8-byte aligned");
  TSPMI_CHECK_ARG(xy_all != nullptr || ixy_all != nullptr || coord_min_n > max_n,
                  "tspmi_batch_fused_generation: coordinates needed for n >= coord_min_n");
  TSPMI_CHECK_ARG(D_all != nullptr || coord_min_n <= 4, "tspmi_batch_fused_generation: a matrix is needed below "
                  "coord_min_n");
  // child: p1, p2, inv1, inv2, child, jump (6n); elite 2-opt: tour, edges, positions (3n + 1) and
  // with int32 coordinates their copy (cand_lds_words + 2n); survive: 2P keys as two 32-bit halves
  // and the P survivor sources (5P)
  size_t words = (size_t)6 * max_n;
  if ((size_t)cand_lds_words(max_n) + 2 * (size_t)max_n > words) words = (size_t)cand_lds_words(max_n) + 2 * max_n;
  if ((size_t)5 * max_p > words) words = (size_t)5 * max_p;
  const size_t lds = words * sizeof(int32_t);
  TSPMI_CHECK_ARG(max_n >= 2 && lds + 2048 <= (size_t)kMaxLds, "tspmi_batch_fused_generation: max_n=%d exceeds LDS",
                  max_n);
  hipStream_t s = (hipStream_t)stream;
  const int2* ct = reinterpret_cast<const int2*>(table_all);
  int rc = TSPMI_OK;
#define TSPMI_BATCH_FUSED(BL, KPV, KLV)                                                                          \
  rc = fused_launch<BL, KLV>(grid, lds, s, desc, B, tasks, ntasks, counters, children_per_task, popA, keysA, off,    \
                             order, popB, keysB, D_all, xy_all, ixy_all, ct, tour_cand, k_tour, cut_lo, cut_hi, probs,   \
                             pos1,                                                                                     \
                             pos2, rate, elite, k, max_iter, coord_min_n, stats, best, best_count)
  if (block == 256) {
    TSPMI_CAND_BY_K(TSPMI_BATCH_FUSED, 256, k)
  } else if (block == 512) {
    TSPMI_CAND_BY_K(TSPMI_BATCH_FUSED, 512, k)
  } else if (block == 1024) {
    TSPMI_CAND_BY_K(TSPMI_BATCH_FUSED, 1024, k)
  } else
    TSPMI_CHECK_ARG(false, "tspmi_batch_fused_generation: block %d not in {256, 512, 1024}", block);
  return rc;
}
