/*
 * tspmi — C ABI of the MI355X-native memetic-TSP hot path (libtspmi.so, gfx950).
 *
 * Conventions
 *   - Every entry point returns an int status: TSPMI_OK (0) or a negative code; the text of
 *     the last failure on the calling thread is returned by tspmi_last_error().
 *   - All array arguments are caller-owned DEVICE pointers (HIP/ROCm global memory),
 *     row-major, contiguous.  Sizes are plain ints.  `stream` is a hipStream_t passed as
 *     void* (NULL = the default stream); every call only enqueues work on that stream and
 *     never synchronises the host, except tspmi_device_count().
 *   - The library never allocates device memory behind the caller's back except the
 *     per-call scratch that an entry point documents.
 *   - Tours are permutations of 0..n-1 stored as int32 (P, n); distance matrices are int32
 *     (n, ld) with 0 <= D < 2^29 so every 2-opt delta fits an int32.
 *
 * Each entry point cites the reference (lucasgaldinos/tsp_gpu_clean) interface it replaces.
 */
#ifndef TSPMI_H
#define TSPMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSPMI_OK 0
#define TSPMI_EINVAL (-1) /* bad argument (shape, mode, null pointer, size limit) */
#define TSPMI_EHIP (-2)   /* HIP runtime error (launch, device) */
#define TSPMI_ENODEV (-3) /* no HIP device */

#define TSPMI_MODE_BEST 0
#define TSPMI_MODE_FIRST 1
#define TSPMI_MODE_MULTI 2

/* Library identification and device probe. */
int tspmi_version(void);
const char* tspmi_last_error(void);
int tspmi_device_count(int* count);

/*
 * Calibration probe of random 4-byte gathers inside 4 KB rows (the distance-gather shape of the
 * full-neighbourhood 2-opt scan): `blocks` workgroups of 256 threads, each thread issuing
 * 8 * iters loads at pseudo-random offsets of row (block % nrows) of `rows` (nrows x 1024 int32).
 * `out` (one int32) is written only if an impossible checksum appears.  Used to price the
 * on-chip gather roofline of the 2-opt scan.
 * No reference counterpart (the reference measures no hardware ceilings).
 */
int tspmi_probe_gather(const int32_t* rows, int nrows, int blocks, int iters, int32_t* out, void* stream);

/*
 * EUC_2D TSPLIB nint distances: D[u,v] = (int)(sqrt(dx^2 + dy^2) + 0.5) in FP64 with a
 * correctly rounded sqrt, diagonal 0.  xy is (n, 2) float64.  Writes max(D) to *max_out
 * (device int32, may be NULL).
 * Replaces: src/distances/matrix.py:214-273 (compute_distance_matrix, EUC_2D branch :84-86)
 *           and src/distances/kernels/pairwise_device.cuh:53-60 (gpu_nint / gpu_euc_2d).
 */
int tspmi_dist_euc2d(const double* xy, int n, int32_t* D, int ld, int32_t* max_out, void* stream);

/*
 * EUC_2D matrix from int32 coordinates (n, 2) by the integer form of the same distance that the
 * full-neighbourhood scan uses for integer instances: equal to tspmi_dist_euc2d on the same
 * (integer-valued) coordinates whenever every pair differs by at most 2^21 per axis.  A pair
 * beyond that gets -1 and is counted in *bad_out (device int32, may be NULL).  xy 8-byte aligned.
 * No reference counterpart: it exposes the integer form so that it can be checked entry by entry
 * against the FP64 one (src/distances/matrix.py:84-86).
 */
int tspmi_dist_euc2d_i32(const int32_t* xy, int n, int32_t* D, int ld, int32_t* bad_out, void* stream);

/*
 * Distance matrix of another TSPLIB 2D metric: metric 0 EUC_2D, 1 CEIL_2D, 2 MAN_2D, 3 MAX_2D,
 * 4 ATT, 5 GEO (xy as DDD.MM latitude/longitude), each with the operations of the NumPy
 * expression it replaces; diagonal 0; max_out as for tspmi_dist_euc2d.  n <= 65535.
 * Replaces: src/distances/matrix.py:77-131 and pairwise.py:62-102 (2D metrics).
 */
int tspmi_dist_metric(const double* xy, int n, int metric, int32_t* D, int ld, int32_t* max_out, void* stream);

/*
 * k-nearest-neighbour candidate table: out[u, r] = r-th nearest v != u in (D[u,v], v)
 * order, k_eff = min(k, n-1) columns.  `ties_out` (device int32, may be NULL) receives the
 * number of rows whose k-th and (k+1)-th distances tie (where NumPy's argpartition choice
 * is implementation-defined).
 * Replaces: src/distances/matrix.py:140-211 (compute_candidate_list).
 */
int tspmi_knn(const int32_t* D, int n, int ld, int k, int32_t* out, int32_t* ties_out, void* stream);

/*
 * tspmi_knn with a per-row flag instead of the count: tie_flags[u] (device int32, n entries) is
 * 1 where row u's k-th and (k+1)-th distances tie, else 0.  Those are the rows on which the
 * reference's np.argpartition picks among equal distances by its own rule; the host rebuilds
 * exactly those rows with that rule (ProblemContext.get_gpu_candidate_list).
 * Replaces: src/distances/matrix.py:140-211 (compute_candidate_list), with the tie rows of
 *           :205-209 left to the caller.
 */
int tspmi_knn_flagged(const int32_t* D, int n, int ld, int k, int32_t* out, int32_t* tie_flags, void* stream);

/*
 * Closed-tour cost: out[r] = (double) sum_x D[t_x, t_{x+1}] (+ D[t_{n-1}, t_0] if closed).
 * Replaces: src/algorithms/strategies/evaluation/tour_cost.py:17-65 and
 *           evaluation/kernels/tour_cost.cu:21 (tour_cost_kernel).
 */
int tspmi_tour_cost(const int32_t* tours, int P, int n, const int32_t* D, int ld, int closed,
                    double* out, void* stream);

/*
 * Population-batched 2-opt on cyclic TSP tours, refined in place.
 *   mode: TSPMI_MODE_BEST (steepest, ties -> lowest (i,j) / (i,r) scan index),
 *         TSPMI_MODE_FIRST (first improving pair in scan order) or
 *         TSPMI_MODE_MULTI (every improving pair of the pass, steepest first with ties in scan
 *         order, applied unless its window [i, j] meets an already accepted one; two_opt.py:491-563;
 *         needs about 30 n bytes of LDS, so n <= ~5000).
 *   cand: (n, k) candidate table or NULL/k=0 for the full neighbourhood.
 *   max_iter: passes per tour; < 0 runs each tour to its local optimum, which requires a
 *             symmetric matrix (on an asymmetric one the delta formula does not make the cost
 *             decrease and the moves can cycle forever: pass a bound there).
 *   stats: device int64[3] accumulated with {deltas evaluated, moves applied, passes}
 *          (may be NULL).
 * Replaces: src/algorithms/strategies/local_search/two_opt.py:831-899 (_improve_grid) with
 *           kernels/two_opt_batch.cu:51 (two_opt_batch_kernel), one launch per call instead of
 *           one launch per pass, and the CPU acceptance rules of two_opt.py:386-489.
 */
int tspmi_two_opt(int32_t* tours, int P, int n, const int32_t* D, int ld, const int32_t* cand, int k,
                  int mode, int max_iter, int64_t* stats, void* stream);

/*
 * tspmi_two_opt(mode = BEST, no candidates) for EUC_2D instances with n <= 2050: D must be the
 * matrix tspmi_dist_euc2d built from xy (float64 (n, 2), 16-byte aligned).  Thirteen rows in
 * sixteen of the scan compute their distances from the coordinates, the others gather them from
 * D.  When every coordinate is an integer and each axis spans at most 2^20 (the kernel checks),
 * the computed distances are lower bounds at most 1 below the matrix entries and each pass
 * settles its best move by re-walking, with D, every diagonal whose bound comes within 2 of the
 * smallest; otherwise they are the matrix's own FP64 values.  Either way the moves, tours and
 * counters are those of tspmi_two_opt.
 * Replaces: two_opt.py:831-899 (full neighbourhood, best) on EUC_2D problems.
 */
int tspmi_two_opt_full_euc2d(int32_t* tours, int P, int n, const int32_t* D, int ld, const double* xy, int max_iter,
                        int64_t* stats, void* stream);

/*
 * tspmi_two_opt_full_euc2d on the instance renumbered in Hilbert-curve order of its coordinates:
 * the nodes' new numbers follow a space-filling curve, so the scan's distance gathers of
 * neighbouring tour positions fall into fewer cache lines once the tours are partly optimised.
 * In one call, on `stream`: the order and its inverse, the renumbered matrix and coordinates
 * (into ws), the tours renumbered in place, the scan, and the tours numbered back.  Renumbering
 * changes no delta and no position, so the moves, tours and counters are tspmi_two_opt_full_euc2d's.
 *   ws: device workspace of at least tspmi_two_opt_full_euc2d_ws_bytes(n) bytes, 256-byte aligned
 *       (overwritten; contents need not be preserved between calls).
 * Replaces: as tspmi_two_opt_full_euc2d.
 */
int tspmi_two_opt_full_euc2d_ordered(int32_t* tours, int P, int n, const int32_t* D, int ld, const double* xy,
                                     int max_iter, int64_t* stats, void* ws, size_t ws_bytes, void* stream);
/* Bytes of workspace tspmi_two_opt_full_euc2d_ordered needs for n cities (n * n * 4 and a little). */
size_t tspmi_two_opt_full_euc2d_ws_bytes(int n);

/*
 * Packed candidate table for tspmi_two_opt_packed: out is (n, KP) pairs of int32
 * {N[u, r], D[u, N[u, r]]} (pairs {-1, 0} for k <= r < KP) with KP = 8, 16 or 32 — the smallest of
 * those >= k.  Build once per (instance, k); 1 <= k <= 32.
 * Replaces: the per-pass candidate lookups of src/algorithms/strategies/local_search/
 *           kernels/two_opt_batch.cu:122-157 (candidate scan), with the first distance of every
 *           candidate pair precomputed.
 */
int tspmi_candidate_table(const int32_t* cand, int n, int k, const int32_t* D, int ld, int32_t* out, void* stream);

/*
 * Steepest (best-mode) candidate-list 2-opt on a packed table from tspmi_candidate_table: the
 * same pairs, order, tie-break and moves as tspmi_two_opt(mode = BEST, cand, k), faster.
 * Replaces: src/algorithms/strategies/local_search/two_opt.py:831-899 in candidate mode.
 */
int tspmi_two_opt_packed(int32_t* tours, int P, int n, const int32_t* D, int ld, const int32_t* table, int k,
                         int max_iter, int64_t* stats, void* stream);

/*
 * tspmi_two_opt_packed that also writes cost[p] = closed cost of tour p after the search (P
 * float64; NULL for none), the value tspmi_tour_cost(closed = 1) gives, without a second launch.
 * Replaces: the local-search-then-evaluate pair of the elite step, pipeline_steps.py:166-183
 *           (EliteOnlyLocalSearchStep: ls.improve_batch, then evaluation.evaluate_batch).
 */
int tspmi_two_opt_packed_cost(int32_t* tours, int P, int n, const int32_t* D, int ld, const int32_t* table, int k,
                              int max_iter, int64_t* stats, double* cost, void* stream);

/*
 * tspmi_two_opt_packed_cost for a few tours (the GA's elite set): when P tours leave CUs idle,
 * each tour is searched by a cluster of up to 16 workgroups that share its candidate pairs and
 * meet once per pass at a bounded barrier in `sync`; otherwise this is tspmi_two_opt_packed_cost.
 * Same tours, moves, counters and costs either way.  best (device int32[1], may be NULL; needs
 * cost) receives the index of the first minimum of the P costs, as tspmi_rank_select(cost, P, 1).
 *   sync: device int64[528 * P], zeroed before the first call; every call leaves it ready for the
 *         next (its flags and counters cleared).  Word 528 p + 512 is set when tour p's cluster
 *         was not co-resident and its barrier timed out (that tour and its cost are then left
 *         unchanged): check it after the work completes and clear it before the next call (a
 *         word still set makes that tour's cluster leave at once, its tour and cost untouched).
 *         Calls that overlap on one `sync` race.
 * Replaces: the elite local search of pipeline_steps.py:166-183 (EliteOnlyLocalSearchStep) with
 *           two_opt.py:831-899 in candidate mode on the elite rows.
 */
int tspmi_two_opt_elite(int32_t* tours, int P, int n, const int32_t* D, int ld, const int32_t* table, int k,
                        int max_iter, int64_t* stats, double* cost, int32_t* best, uint64_t* sync, void* stream);

/*
 * tspmi_two_opt_elite for EUC_2D instances with int32 coordinates ixy ((n, 2), 8-byte aligned,
 * each axis spanning at most 2^20): the distances of the scan and of the refreshed edges are
 * computed from the coordinates held in LDS (the integer form, equal to the matrix entries), so
 * no matrix is read and a pass touches no global memory but the cluster's barrier words.  Where
 * no cluster applies, this is tspmi_two_opt_packed_euc2d_i32 with costs.  Same tours, moves,
 * counters, costs and best index as tspmi_two_opt_elite on the matrix of the same coordinates.
 * Replaces: as tspmi_two_opt_elite.
 */
int tspmi_two_opt_elite_xy(int32_t* tours, int P, int n, const int32_t* ixy, const int32_t* table, int k,
                           int max_iter, int64_t* stats, double* cost, int32_t* best, uint64_t* sync, void* stream);

/*
 * tspmi_two_opt_elite without a caller workspace, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the cluster barrier
 * words live in a static device array that the library holds on every GPU it runs on until the
 * process ends (TSPMI_ELITE_STATIC_TOURS x 528 int64 = 540,672 bytes per GPU), so P may not
 * exceed TSPMI_ELITE_STATIC_TOURS (TSPMI_EINVAL otherwise).  The call allocates nothing.  Each
 * call first clears the P tours' words with a kernel, so nothing of an earlier call carries over;
 * failed (device int32[1], may be NULL) receives the number of tours whose cluster barrier timed
 * out (those tours and their costs are left unchanged, as in tspmi_two_opt_elite).  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU race.  Same tours, moves,
 * counters, costs and best index as tspmi_two_opt_elite.
 * No reference counterpart: the reference allocates its workspaces per call.
 */
#define TSPMI_ELITE_STATIC_TOURS 128
int tspmi_two_opt_elite_noalloc(int32_t* tours, int P, int n, const int32_t* D, int ld, const int32_t* table, int k,
                                int max_iter, int64_t* stats, double* cost, int32_t* best, int32_t* failed,
                                void* stream);

/*
 * EUC_2D variants that compute every distance from the coordinates xy (n, 2) float64 (16-byte
 * aligned) with the FP64 nint of tspmi_dist_euc2d instead of reading a matrix: same tours, moves
 * and counters as tspmi_two_opt(mode = BEST, full neighbourhood) and tspmi_two_opt_packed on the
 * matrix those coordinates produce.  The full-neighbourhood form needs n <= 8194.
 * Replaces: src/algorithms/strategies/local_search/two_opt.py:831-899 (best mode) for EUC_2D
 *           problems, reading coordinates where the reference reads context.get_gpu_distances().
 */
int tspmi_two_opt_euc2d(int32_t* tours, int P, int n, const double* xy, int max_iter, int64_t* stats, void* stream);
int tspmi_two_opt_packed_euc2d(int32_t* tours, int P, int n, const double* xy, const int32_t* table, int k,
                               int max_iter, int64_t* stats, void* stream);
/* tspmi_two_opt_packed_euc2d with the final closed costs (as tspmi_two_opt_packed_cost);
 * costs need n >= 4 and max_iter != 0. */
int tspmi_two_opt_packed_euc2d_cost(int32_t* tours, int P, int n, const double* xy, const int32_t* table, int k,
                                    int max_iter, int64_t* stats, double* cost, void* stream);

/*
 * tspmi_two_opt_packed_euc2d_cost that also logs each tour's progress: pass_log (device int64,
 * P x (log_moves + 1) x 2) receives per tour {cost, wall clock} at the start and after each of its
 * first log_moves moves (entries past a tour's last logged move are left as they were).  The cost
 * is exact (start cost plus the applied deltas); the clock is the device's constant-rate counter,
 * in ticks of tspmi_wall_clock_khz() kHz, common to all CUs.  With it the population's best cost
 * is known at the resolution of single passes from one launch, without host synchronisation.
 * Same tours, counters and costs as tspmi_two_opt_packed_euc2d_cost.
 * Replaces: two_opt.py:831-899 in candidate mode (the per-pass best-cost trace has no counterpart:
 *           the reference's convergence history is kept per generation, genetic_algorithm.py:951-964).
 */
int tspmi_two_opt_packed_euc2d_log(int32_t* tours, int P, int n, const double* xy, const int32_t* table, int k,
                                   int max_iter, int64_t* stats, double* cost, int64_t* pass_log, int log_moves,
                                   void* stream);

/*
 * tspmi_two_opt_packed_euc2d(_cost, _log) from int32 coordinates ixy (n, 2), 8-byte aligned, by the
 * integer form of the EUC_2D distance: the caller guarantees integer-valued coordinates whose
 * per-axis span is at most 2^20, where the integer form equals the FP64 one (else the distances
 * are wrong).  cost and pass_log may be NULL (log_moves is then ignored).  Same tours, moves,
 * counters, costs and log costs as the float64 entry points on the same coordinates; about half the
 * VALU work and half the coordinate bytes per distance.
 * Replaces: two_opt.py:831-899 in candidate mode on EUC_2D problems with integer coordinates.
 */
int tspmi_two_opt_packed_euc2d_i32(int32_t* tours, int P, int n, const int32_t* ixy, const int32_t* table, int k,
                                   int max_iter, int64_t* stats, double* cost, int64_t* pass_log, int log_moves,
                                   void* stream);

/*
 * tspmi_two_opt_packed_euc2d_i32 on the instance renumbered in Hilbert-curve order of its
 * coordinates (as tspmi_two_opt_full_euc2d_ordered): in one call, the order (tspmi_rank_select of
 * the nodes' curve indices), the renumbered coordinates and candidate table (into ws), the tours
 * renumbered in place, the scan, and the tours numbered back.  Same moves, tours, counters and
 * costs; the pass log's clock starts with the scan.  Shapes the renumbering does not apply to,
 * and tours small enough for the scan's node records in LDS (n up to about 5700, where it gains
 * nothing), go to tspmi_two_opt_packed_euc2d_i32 unchanged.
 *   ws: device workspace of tspmi_two_opt_packed_euc2d_ws_bytes(n, k) bytes, 256-byte aligned.
 * Replaces: as tspmi_two_opt_packed_euc2d_i32.
 */
int tspmi_two_opt_packed_euc2d_i32_ordered(int32_t* tours, int P, int n, const int32_t* ixy, const int32_t* table,
                                           int k, int max_iter, int64_t* stats, double* cost, int64_t* pass_log,
                                           int log_moves, void* ws, size_t ws_bytes, void* stream);
/* Bytes of workspace tspmi_two_opt_packed_euc2d_i32_ordered needs (0 for k outside [1, 32]). */
size_t tspmi_two_opt_packed_euc2d_ws_bytes(int n, int k);

/*
 * Best-improvement 2-opt for instances too large for the LDS-resident kernels (candidate scan
 * above n ~ 13,500, full neighbourhood above n = 16,386): each tour's positions, edge costs and
 * (candidate mode) inverse positions live in the caller's workspace ws, device int32[P * (3n + 1)],
 * scratch whose contents need not be initialised and are dead after the call.  table (packed, as
 * tspmi_two_opt_packed) selects the candidate scan with k candidates, NULL the full neighbourhood.
 * Distances come from xy (EUC_2D coordinates, float64 (n, 2), computed as tspmi_dist_euc2d does)
 * when xy is given, else from D (n * ld < 2^30).  cost (nullable) receives each final closed cost.
 * Same pairs, order, tie-break, moves and counters as tspmi_two_opt / tspmi_two_opt_packed.
 * Replaces: two_opt.py:831-899 above the shared-memory tile limit (two_opt.py:75-94, 707-709;
 *           kernels/two_opt_batch.cu:192-268, the windowed scan).
 */
int tspmi_two_opt_large(int32_t* tours, int P, int n, const int32_t* D, int ld, const double* xy, const int32_t* table,
                        int k, int max_iter, int64_t* stats, double* cost, int32_t* ws, void* stream);

/* Rate of the device's constant wall clock (the clock of tspmi_two_opt_packed_euc2d_log), in kHz.
 * No reference counterpart. */
int tspmi_wall_clock_khz(int* khz);

/*
 * Tournament selection: out[r] = pop[w_r] with w_r = cand[r, argmin_c fitness[cand[r, c]]]
 * (first minimum).  cand is (num_parents, k) int32 drawn on the host.
 * Replaces: src/algorithms/strategies/selection/tournament_selection.py:78-128 and
 *           selection/kernels/tournament_selection.cu:11.
 */
int tspmi_tournament(const int32_t* pop, int P, int n, const double* fitness, const int32_t* cand, int k,
                     int num_parents, int32_t* out, int32_t* winners_out, void* stream);

/*
 * Partially mapped crossover of pairs (2q, 2q+1) with shared cuts [lo[q], hi[q]]
 * (lo <= hi); an odd last row is copied.
 * Replaces: src/algorithms/strategies/crossover/pmx_crossover.py:49-171 and
 *           crossover/kernels/pmx_crossover.cu:19.
 */
int tspmi_pmx(const int32_t* parents, int P, int n, const int32_t* cut_lo, const int32_t* cut_hi,
              int32_t* out, void* stream);

/*
 * Inversion mutation in place: rows with probs[r] < rate reverse [min(p1,p2), max(p1,p2)].
 * Replaces: src/algorithms/strategies/mutation/inversion_mutation.py:65-214 and
 *           mutation/kernels/inversion_mutation.cu:11.
 */
int tspmi_inversion(int32_t* routes, int P, int n, const double* probs, const int32_t* pos1,
                    const int32_t* pos2, double rate, void* stream);

/*
 * Order crossover (OX) of pairs (2q, 2q+1) with shared cuts [lo[q], hi[q]]: child keeps its own
 * parent's segment and takes the other parent's remaining genes in that parent's order from
 * position hi + 1 on (wrapping); an odd last row is copied.
 * Replaces: src/algorithms/strategies/crossover/order_crossover.py:49-176 and
 *           crossover/kernels/order_crossover.cu.
 */
int tspmi_ox(const int32_t* parents, int P, int n, const int32_t* cut_lo, const int32_t* cut_hi, int32_t* out,
             void* stream);

/*
 * Swap mutation in place: rows with probs[r] < rate exchange positions pos1[r] and pos2[r].
 * Replaces: src/algorithms/strategies/mutation/swap_mutation.py:84-105 (uniform positions) and
 *           mutation/kernels/swap_mutation.cu.
 */
int tspmi_swap(int32_t* routes, int P, int n, const double* probs, const int32_t* pos1, const int32_t* pos2,
               double rate, void* stream);

/*
 * Host (no device work): G generations of the draws tspmi_generation consumes, in the order the
 * reference strategies take them from NumPy's Generator(PCG64), bit-identical to those calls:
 * per generation cand (P, k) = integers(0, P); cut pairs: P/2 first points integers(0, n), then
 * P/2 second points integers(0, n - 1) shifted past the first, stored as (min, max) in cut_lo /
 * cut_hi (rows of max(1, P/2)); probs (P) = random(); pos1 (P) = integers(0, n); pos2 (P) =
 * integers(0, n - 1) shifted past pos1.  `state` is {state_hi, state_lo, inc_hi, inc_lo,
 * has_uint32, uinteger} of the generator (NumPy's bit_generator.state) and is advanced in place;
 * `states_out` (nullable, G x 6) receives it after each generation.
 * Replaces: the per-generation draws of tournament_selection.py:113, pmx_crossover.py:155 with
 *           route_helpers.py:74-108, and inversion_mutation.py:84-92.
 */
int tspmi_draw_generations(uint64_t* state, int G, int P, int n, int k, int32_t* cand, int32_t* cut_lo,
                           int32_t* cut_hi, double* probs, int32_t* pos1, int32_t* pos2, uint64_t* states_out);

/*
 * Fused generation: tournament -> PMX -> inversion -> evaluate, one workgroup per child row.
 * Inputs are the host-drawn buffers of one generation: tour_cand (P, k), cut_lo/cut_hi (P/2),
 * probs/pos1/pos2 (P).  Writes offspring (P, n) and their costs off_fit (P).
 * Replaces: the SelectionStep -> CrossoverStep -> MutationStep -> EvaluationStep sequence of
 *           src/algorithms/metaheuristics/genetic_algorithm.py:120-213.
 */
int tspmi_generation(const int32_t* pop, const double* fitness, int P, int n, const int32_t* D, int ld,
                     const int32_t* tour_cand, int k, const int32_t* cut_lo, const int32_t* cut_hi,
                     const double* probs, const int32_t* pos1, const int32_t* pos2, double rate,
                     int32_t* offspring, double* off_fit, void* stream);

/*
 * Stable rank selection: for the M keys, rank[i] = #{j : key[j] < key[i] or
 * (key[j] == key[i] and j < i)}; writes order[rank[i]] = i for rank[i] < keep.
 * With keys = concat(old_fit, new_fit) this is the (mu + lambda) survivor order of
 * genetic_algorithm.py:618-706 (lexsort :689); with keys = fitness and keep = 16 it is the
 * elite set of pipeline_steps.py:153-163.  Keys compare as IEEE totally-ordered bits with -0.0
 * equal to +0.0 (NaN sorts after +inf), so the order is defined for every input.  M <= 2^24.
 */
int tspmi_rank_select(const double* keys, int M, int keep, int32_t* order, void* stream);

/*
 * Row gather of two sources: out[d] = (src[d] < P0 ? A[src[d]] : B[src[d] - P0]) for d < rows,
 * and out_fit[d] likewise from fitA/fitB (out_fit may be NULL).
 * Replaces: the combined_pop[best_indices] gather of genetic_algorithm.py:628,706.
 */
int tspmi_gather_rows(const int32_t* A, const double* fitA, int P0, const int32_t* B, const double* fitB,
                      int n, const int32_t* src, int rows, int32_t* out, double* out_fit, void* stream);

/*
 * (mu + lambda) survival in one launch: tspmi_rank_select over keys (M float64: the fitness of
 * A's P0 rows, then of B's M - P0 rows), then the survivors' rows in stable rank order,
 * out[r] = row order[r] of A (order[r] < P0) or of B, with out_fit[r] = keys[order[r]] and
 * order[r] itself (out_fit and order may be NULL), for r < keep.  Same result as tspmi_rank_select
 * followed by tspmi_gather_rows, without the second launch.  out must not overlap A or B.
 * Replaces: _select_survivors_assessed, genetic_algorithm.py:618-706 (lexsort, then the
 *           combined_pop[best_indices] gather).
 */
int tspmi_survive(const double* keys, int M, int keep, const int32_t* A, int P0, const int32_t* B, int n,
                  int32_t* out, double* out_fit, int32_t* order, void* stream);

/*
 * Row scatter: dst[idx[r]] = src[r] (and dst_fit[idx[r]] = src_fit[r] when non-NULL).
 * Replaces: the elite write-back of pipeline_steps.py:183-184.
 */
int tspmi_scatter_rows(const int32_t* src, const double* src_fit, const int32_t* idx, int rows, int n,
                       int32_t* dst, double* dst_fit, void* stream);

/*
 * SplitMix64 Fisher-Yates permutations of `customers` (length n), one stream per row seed.
 * Replaces: src/algorithms/strategies/initialization/random_init.py:24-118 and
 *           initialization/kernels/random_init.cu:9.
 */
int tspmi_random_init(const uint64_t* seeds, int P, const int32_t* customers, int n, int32_t* out,
                      void* stream);

/*
 * GRASP nearest-neighbour tours over a (n, k) candidate table with SplitMix64 draws.
 * `work` is caller-owned device scratch of 2 * P * n int32.
 * Replaces: src/algorithms/strategies/initialization/nearest_neighbor_init.py:31-238 and
 *           initialization/kernels/nearest_neighbor_init.cu:10.
 */
int tspmi_nn_init(const uint64_t* seeds, int P, const int32_t* customers, int n, const int32_t* cand,
                  int k, int32_t* work, int32_t* out, void* stream);

/*
 * Greedy nearest-neighbour tours over cities 0..n-1: row r starts at starts[r] and always moves to
 * the nearest unvisited city (lowest index on equal distances).  One workgroup per row.
 * No counterpart in the reference (its nearest-neighbour initialiser is the randomised one
 * above); used to build common start tours when comparing candidate-list schemes.
 */
int tspmi_nn_greedy_init(const int32_t* starts, int P, int n, const int32_t* D, int ld, int32_t* out,
                         void* stream);

/* ---------------------------------------------------------------------------------------------
 * Uneven multi-instance batch (many instances of different n, each with its own population).
 *
 * `desc` is a device int64 table with TSPMI_BATCH_FIELDS records per instance:
 *   [0] n, [1] P, [2] xy offset (float64 elements), [3] D offset (int32; D is n x n),
 *   [4] packed candidate table offset (int2 pairs; n x 32), [5] k-NN offset (int32; n x k),
 *   [6] population offset (int32; P x n), [7] row offset (sum of P over earlier instances),
 *   [8] pair offset (sum of P/2 over earlier instances).
 * Key buffers hold 2P float64 per instance at 2 * row offset: parents' fitness then offspring's.
 * Host-drawn random buffers of one generation are concatenated in instance order: tour_cand
 * (rows x k), cut_lo/cut_hi (pairs), probs/pos1/pos2 (rows).
 * Work lists are flat (item_inst, item_idx) int32 arrays; order them longest-first.  Every stage
 * runs the same device code as the single-instance entry points, so each instance evolves exactly
 * as it would alone (genetic_algorithm.py:789-965 per instance).
 * ------------------------------------------------------------------------------------------- */
#define TSPMI_BATCH_FIELDS 9

/* Distance rows: items are (instance, row u).  Replaces matrix.py:214-273 per instance. */
int tspmi_batch_dist_euc2d(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems,
                           const double* xy_all, int32_t* D_all, void* stream);

/* k-NN rows (k <= 32) and packed candidate rows: items are (instance, row u).  Replaces
 * matrix.py:140-211 per instance. */
int tspmi_batch_knn(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems, int k,
                    const int32_t* D_all, int32_t* cand_all, int32_t* table_all, void* stream);

/* SplitMix64 random permutations of 0..n-1: items are (instance, row); seeds indexed by row
 * offset + row.  Replaces random_init.py:24-118 per instance. */
int tspmi_batch_random_init(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems,
                            const uint64_t* seeds, int32_t* pop, void* stream);

/* Tour costs of listed (instance, row) items into key slot 2 * row offset + row; with xy_all or
 * ixy_all (either may be NULL) the EUC_2D distances are computed from the coordinates instead of
 * read from D_all (same values).  ixy_all holds the same coordinates as int32 at the same offsets
 * (descriptor field [2]); when given, it is used, by the integer distance form, and every
 * instance's coordinates must be integers spanning at most 2^20 per axis.
 * Replaces tour_cost.py:17-65 per instance. */
int tspmi_batch_cost(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems,
                     const int32_t* pop, const int32_t* D_all, const double* xy_all, const int32_t* ixy_all,
                     double* keys, void* stream);

/* Fused generation of every instance: items are (instance, child row r); block in
 * {128, 256, 512, 1024}; grid <= 0 means one workgroup per item, else a persistent grid; child
 * costs from xy_all or ixy_all when given (as tspmi_batch_cost).
 * Replaces genetic_algorithm.py:120-213 per instance. */
int tspmi_batch_generation(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems,
                           int max_n, const int32_t* popA, double* keysA, const int32_t* D_all, const double* xy_all,
                           const int32_t* ixy_all, const int32_t* tour_cand, int k, const int32_t* cut_lo, const int32_t* cut_hi,
                           const double* probs, const int32_t* pos1, const int32_t* pos2, double rate, int32_t* off,
                           int block, int grid, void* stream);

/* Stable (mu + lambda) survivor order of every instance (2P <= 4096) into order (P per instance,
 * at row offset).  Replaces genetic_algorithm.py:618-706 per instance. */
int tspmi_batch_survival(const int64_t* desc, int B, const double* keysA, int32_t* order, void* stream);

/* Survivor rows and keys into popB/keysB in rank order: items are (instance, destination row). */
int tspmi_batch_gather(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems,
                       const int32_t* popA, const double* keysA, const int32_t* off, const int32_t* order,
                       int32_t* popB, double* keysB, void* stream);

/* Best-mode candidate 2-opt of listed (instance, row) tours in place, block in {256, 512, 1024},
 * grid <= 0 one workgroup per item, else persistent.  Replaces pipeline_steps.py:153-195 +
 * two_opt.py:831-899 per instance. */
int tspmi_batch_two_opt_packed(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx, int nitems,
                               int max_n, int32_t* pop, const int32_t* D_all, const int32_t* table_all, int k,
                               int max_iter, int block, int grid, int64_t* stats, void* stream);

/* As tspmi_batch_two_opt_packed; instances with n >= coord_min_n compute EUC_2D distances from
 * xy_all (descriptor field [2]) instead of reading their matrix (same tours); from ixy_all, by the
 * integer form, when it is given (as tspmi_batch_cost). */
int tspmi_batch_two_opt_packed_mixed(const int64_t* desc, const int32_t* item_inst, const int32_t* item_idx,
                                     int nitems, int max_n, int32_t* pop, const int32_t* D_all, const double* xy_all,
                                     const int32_t* ixy_all, const int32_t* table_all, int k, int max_iter, int block,
                                     int grid, int coord_min_n, int64_t* stats, void* stream);

/* First minimum among the first `count` keys of every instance.  Replaces genetic_algorithm.py:708-739. */
int tspmi_batch_best(const int64_t* desc, int B, const double* keys, int count, int32_t* best, void* stream);

/* One whole generation of every instance in ONE persistent launch: the fused child generation,
 * the stable (mu + lambda) survival, the survivors' copy into popB/keysB and the elite candidate
 * 2-opt with its costs (plus, when `best` is given, the first minimum of each instance's first
 * best_count keys), with the same results as tspmi_batch_generation + tspmi_batch_survival +
 * tspmi_batch_gather + tspmi_batch_two_opt_packed_mixed + tspmi_batch_cost (+ tspmi_batch_best).
 * `tasks` is a device list of ntasks uint32 task words, kind << 30 | instance << 12 | index, in
 * this order: every child task (0, i, g) for g < ceil(P / children_per_task) (children
 * g * children_per_task onwards; children_per_task even and >= 2), then every survive task
 * (1, i, 0), then every elite row (2, i, d) for d < min(elite, P); instances longest first within
 * each kind.  `counters` is a device int32 array of 4 + 3 B words, zero before the first call and
 * owned by one batch: word 2 is set to 1 if a dependency wait of a call ran out (that call's
 * results are then invalid), word 3 counts calls, the others are back at zero after every call,
 * so consecutive calls on one stream and graph replays need no clearing; calls that share
 * `counters` must not overlap.  block in {256, 512, 1024}; grid <= 0 launches as many workgroups
 * as fit on the device at once.  Distances from ixy_all / xy_all as in tspmi_batch_two_opt_packed_mixed.
 * Replaces genetic_algorithm.py:120-296 (selection, crossover, mutation, evaluation, survival
 * steps) and pipeline_steps.py:153-195 (EliteOnlyLocalSearchStep) per instance; the reference
 * has no multi-instance batch. */
int tspmi_batch_fused_generation(const int64_t* desc, int B, const uint32_t* tasks, int ntasks, int32_t* counters,
                                 int children_per_task, int max_n, int max_p, const int32_t* popA, double* keysA,
                                 int32_t* off, int32_t* order, int32_t* popB, double* keysB, const int32_t* D_all,
                                 const double* xy_all, const int32_t* ixy_all, const int32_t* table_all,
                                 const int32_t* tour_cand,
                                 int k_tour, const int32_t* cut_lo, const int32_t* cut_hi, const double* probs,
                                 const int32_t* pos1, const int32_t* pos2, double rate, int elite, int k,
                                 int max_iter, int coord_min_n, int block, int grid, int64_t* stats, int32_t* best,
                                 int best_count, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Learned candidate-edge scorer.  No counterpart exists in the reference (it extends the
 * reference's k-nearest candidate lists of src/distances/matrix.py:140-211 and
 * two_opt.py:117-149); the C ABI follows the same conventions.
 * Parameters: one fp32 master buffer and one bf16 (uint16) buffer of tspmi_scorer_param_count()
 * values: W1 [128][16] (features 12..15 zero), b1 [128], W2 [128][128], b2 [128], w3 [128], b3.
 * ------------------------------------------------------------------------------------------- */
int tspmi_scorer_param_count(void);

/* Features X (n*32, 16) bf16 of (city u, its r-th nearest neighbour), from coordinates, the
 * (n, 32) nearest-neighbour table and the distance matrix; n >= 33. */
int tspmi_scorer_features(const double* xy, int n, const int32_t* nn32, const int32_t* D, int ld, uint16_t* X,
                          void* stream);

/* Labels y (n*32) float: 1 when (u, r-th neighbour) is an edge of `tour` (either direction);
 * work holds n int32. */
int tspmi_scorer_labels(const int32_t* tour, int n, const int32_t* nn32, int32_t* work, float* y, void* stream);

/* Logits z (S) of the MLP for S feature rows (MFMA bf16, fp32 accumulation). */
int tspmi_mlp_forward(const uint16_t* X, int S, const uint16_t* wbf, float* z, void* stream);

/* Gradient of the mean BCE-with-logits over S rows (S % 8 == 0) into grad (param_count + 1
 * floats; the last one is the loss).  work / partials sizes from tspmi_mlp_work_sizes. */
int tspmi_mlp_grad(const uint16_t* X, const float* y, int S, const uint16_t* wbf, uint16_t* work, float* partials,
                   float* grad, void* stream);
int tspmi_mlp_work_sizes(int S, long* work_elems, long* partial_floats);

/* One Adam step on the fp32 master weights; refreshes the bf16 copy. */
int tspmi_adam(float* theta, float* m, float* v, const float* grad, int step, float lr, float beta1, float beta2,
               float eps, float weight_decay, uint16_t* wbf, void* stream);

/* fp32 -> bf16 (round to nearest even). */
int tspmi_to_bf16(const float* src, int count, uint16_t* dst, void* stream);

/* Learned candidate table out (n, k): the k highest-scoring of each city's 32 nearest neighbours
 * (ties to the nearer rank), in decreasing score order; k <= 32. */
int tspmi_scorer_topk(const float* z, const int32_t* nn32, int n, int k, int32_t* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TSPMI_H */
