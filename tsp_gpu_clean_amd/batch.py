"""Uneven multi-instance memetic GA: many TSP instances of different n, each with its own
population and generator, advanced together — either one persistent launch per generation
(``tspmi_batch_fused_generation``: a dependency-ordered task list of children, survival, survivor
copies and elite 2-opt of every instance) or one launch per pipeline stage covering every
instance; whole runs of generations captured in one HIP graph, the form and launch shapes picked
by a small autotuner.

Each instance follows exactly the pipeline of ``DeviceMemeticLoop`` (the reference's
Selection -> PMX -> Inversion -> Evaluation -> Survival -> EliteOnly 2-opt, genetic_algorithm.py
:789-965 with pipeline_steps.py:153-195) with its own PCG64 stream, so a batched run is bit-identical
to running every instance alone.
"""

from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _native
from .context import int_coordinates
from .engine import capture_graph, draw_generations
from .rng import draw_row_seeds, make_rng

FIELDS = 9  # TSPMI_BATCH_FIELDS
KP = 32  # packed candidate row width used by the batch kernels
TASK_CHILD, TASK_SURVIVE, TASK_ELITE = 0, 1, 2
# Instances at least this large compute their distances from coordinates in the elite search (and
# every instance does in the generation and cost kernels).  With 256 instances the matrices (1.5 GB
# at n <= 2000) thrash the caches while each instance's coordinates are at most 32 KB: measured on
# an MI355X at 256 instances, n in [50, 2000], 100 generations, all-coordinates ran 464 gens/s
# against 173 with matrices only (256 / 368 / 449 with the threshold at 1500 / 1000 / 600).
COORD_MIN_N = 0


@dataclass
class BatchDraws:
  tour_cand: object  # (G, rows * k)
  cut_lo: object  # (G, pairs)
  cut_hi: object
  probs: object  # (G, rows)
  pos1: object
  pos2: object


def _cu(a):
  import torch

  return torch.from_numpy(np.ascontiguousarray(a)).cuda()


class BatchedMemeticGA:
  def __init__(self, coords, population_size: int = 64, mutation_rate: float = 0.05, tournament_k: int = 5,
               elite_count: int = 16, ls_max_iterations: int = 10, ls_candidate_k: int = 20, seeds=None,
               integer_distances: bool = True) -> None:
    torch = _native.require_device()
    self.B = len(coords)
    self.ns = np.array([int(np.asarray(c).shape[0]) for c in coords], dtype=np.int64)
    if (self.ns < 4).any():
      raise ValueError("every instance needs n >= 4")
    self.P = int(population_size)
    if 2 * self.P > 4096:
      raise ValueError("batched survival sorts 2P <= 4096 keys per instance")
    self.rate = float(mutation_rate)
    self.k_tour = int(tournament_k)
    self.elite = min(int(elite_count), self.P)
    self.ls_iters = int(ls_max_iterations)
    self.kc = int(ls_candidate_k)
    if not 1 <= self.kc <= 32:
      raise ValueError("ls_candidate_k must lie in [1, 32]")
    seeds = list(range(self.B)) if seeds is None else list(seeds)
    self.rngs = [make_rng(s) for s in seeds]
    self.max_n = int(self.ns.max())
    B, P, ns = self.B, self.P, self.ns
    half = P // 2
    off_xy = np.concatenate([[0], np.cumsum(2 * ns)[:-1]])
    off_d = np.concatenate([[0], np.cumsum(ns * ns)[:-1]])
    off_ct = np.concatenate([[0], np.cumsum(ns * KP)[:-1]])
    off_cand = np.concatenate([[0], np.cumsum(ns * self.kc)[:-1]])
    off_pop = np.concatenate([[0], np.cumsum(P * ns)[:-1]])
    desc = np.zeros((B, FIELDS), dtype=np.int64)
    desc[:, 0] = ns
    desc[:, 1] = P
    desc[:, 2] = off_xy
    desc[:, 3] = off_d
    desc[:, 4] = off_ct
    desc[:, 5] = off_cand
    desc[:, 6] = off_pop
    desc[:, 7] = np.arange(B) * P
    desc[:, 8] = np.arange(B) * half
    self.desc_host = desc
    self.desc = _cu(desc)
    self.rows = B * P
    self.pairs = B * half
    # work lists, longest instance first (LPT)
    order = np.argsort(-ns, kind="stable")

    def items(counts):
      inst = np.concatenate([np.full(int(counts[i]), i, np.int32) for i in order])
      idx = np.concatenate([np.arange(int(counts[i]), dtype=np.int32) for i in order])
      return _cu(inst), _cu(idx), int(inst.size)

    self.w_dist = items(ns)
    self.w_row = items(np.full(B, P))
    self.w_elite = items(np.full(B, self.elite))
    self.set_children_per_task(4)
    self.counters = torch.zeros(4 + 3 * B, dtype=torch.int32, device="cuda")
    xy = np.concatenate([np.ascontiguousarray(c, dtype=np.float64).reshape(-1) for c in coords])
    self.xy = _cu(xy)
    # Integer instances (every one within the 2^20 span) take the integer distance form: the same
    # distances from int32 coordinates at the same offsets, about half the work per distance.
    # ``integer_distances=False`` keeps the FP64 form (same results) for comparison.
    ixy = [int_coordinates(np.asarray(c, dtype=np.float64).reshape(-1, 2)) for c in coords]
    use_int = integer_distances and all(a is not None for a in ixy)
    self.ixy = _cu(np.concatenate([a.reshape(-1) for a in ixy])) if use_int else None
    total_pop = int((P * ns).sum())
    dev = "cuda"
    self.D = torch.empty(int((ns * ns).sum()), dtype=torch.int32, device=dev)
    self.cand = torch.empty(int((ns * self.kc).sum()), dtype=torch.int32, device=dev)
    self.table = torch.empty(int((ns * KP).sum()) * 2, dtype=torch.int32, device=dev)
    self.pop = [torch.empty(total_pop, dtype=torch.int32, device=dev) for _ in range(2)]
    self.off = torch.empty(total_pop, dtype=torch.int32, device=dev)
    self.keys = [torch.zeros(2 * self.rows, dtype=torch.float64, device=dev) for _ in range(2)]
    self.order = torch.empty(self.rows, dtype=torch.int32, device=dev)
    self.stats = torch.zeros(3, dtype=torch.int64, device=dev)
    self.cur = 0
    self.shapes = {"fused": False, "fused_block": 256, "fused_grid": 0, "children_per_task": 4, "gen_block": 256,
                   "gen_grid": 0, "ls_block": 1024, "ls_grid": 0}
    _native.batch_call("tspmi_batch_dist_euc2d", self.desc, *self.w_dist, self.xy, self.D)
    _native.batch_call("tspmi_batch_knn", self.desc, *self.w_dist, self.kc, self.D, self.cand, self.table)

  def set_children_per_task(self, c: int) -> None:
    """Children per child task of the fused generation (even): fewer, longer tasks publish less."""
    if c < 2 or c % 2:
      raise ValueError("children per task must be even and >= 2")
    self.children_per_task = int(c)
    self.tasks_host = fused_task_list(self.ns, self.P, self.elite, self.children_per_task)
    self.tasks = _cu(self.tasks_host.view(np.int32))

  # ---- state ------------------------------------------------------------------------------------
  def init_random(self) -> None:
    """SplitMix64 random permutations from each instance's generator (random_init.py:30-55)."""
    seeds = np.concatenate([draw_row_seeds(r, self.P) for r in self.rngs])
    _native.batch_call("tspmi_batch_random_init", self.desc, *self.w_row, _cu(seeds.view(np.int64)), self.pop[self.cur])
    _native.batch_call("tspmi_batch_cost", self.desc, *self.w_row, self.pop[self.cur], self.D, self.xy, self.ixy,
                       self.keys[self.cur])

  def population(self, i: int):
    """Device view (P, n_i) of instance i's current population."""
    o = int(self.desc_host[i, 6])
    n = int(self.ns[i])
    return self.pop[self.cur][o : o + self.P * n].view(self.P, n)

  def fitness(self, i: int):
    o = 2 * i * self.P
    return self.keys[self.cur][o : o + self.P]

  def distances(self, i: int):
    o = int(self.desc_host[i, 3])
    n = int(self.ns[i])
    return self.D[o : o + n * n].view(n, n)

  # ---- random draws ---------------------------------------------------------------------------
  def draw(self, generations: int) -> BatchDraws:
    """Each instance's draws, generation by generation, in the reference order; concatenated."""
    P, k, G, B = self.P, self.k_tour, generations, self.B
    half = P // 2
    cand = np.empty((G, B, P * k), np.int32)
    lo = np.zeros((G, B, max(1, half)), np.int32)
    hi = np.zeros((G, B, max(1, half)), np.int32)
    probs = np.empty((G, B, P), np.float64)
    p1 = np.empty((G, B, P), np.int32)
    p2 = np.empty((G, B, P), np.int32)
    hp = max(1, half)
    for i, rng in enumerate(self.rngs):
      one = draw_generations(rng, G, P, int(self.ns[i]), k)  # the instance's generator, generation order
      cand[:, i] = one.tour_cand.reshape(G, -1)
      lo[:, i, :hp] = one.cut_lo
      hi[:, i, :hp] = one.cut_hi
      probs[:, i] = one.probs
      p1[:, i] = one.pos1
      p2[:, i] = one.pos2
    return BatchDraws(_cu(cand.reshape(G, -1)), _cu(lo[:, :, :hp].reshape(G, -1)), _cu(hi[:, :, :hp].reshape(G, -1)),
                      _cu(probs.reshape(G, -1)), _cu(p1.reshape(G, -1)), _cu(p2.reshape(G, -1)))

  # ---- one generation ---------------------------------------------------------------------------
  def _generation(self, a, d: BatchDraws, g: int, block: int, grid: int) -> None:
    _native.batch_call("tspmi_batch_generation", self.desc, *self.w_row, self.max_n, self.pop[a], self.keys[a], self.D,
                       self.xy, self.ixy, d.tour_cand[g], self.k_tour, d.cut_lo[g], d.cut_hi[g], d.probs[g], d.pos1[g], d.pos2[g],
                       self.rate, self.off, block, grid)

  def _elite_ls(self, pop, block: int, grid: int) -> None:
    _native.batch_call("tspmi_batch_two_opt_packed_mixed", self.desc, *self.w_elite, self.max_n, pop, self.D, self.xy,
                       self.ixy, self.table, self.kc, self.ls_iters, block, grid, COORD_MIN_N, self.stats)

  def _fused(self, a, b, d: BatchDraws, g: int, best, block: int, grid: int) -> None:
    _native.batch_call("tspmi_batch_fused_generation", self.desc, self.B, self.tasks, int(self.tasks_host.size),
                       self.counters, self.children_per_task, self.max_n, self.P, self.pop[a], self.keys[a], self.off,
                       self.order, self.pop[b], self.keys[b], self.D, self.xy, self.ixy, self.table, d.tour_cand[g],
                       self.k_tour, d.cut_lo[g], d.cut_hi[g], d.probs[g], d.pos1[g], d.pos2[g], self.rate, self.elite,
                       self.kc, self.ls_iters, COORD_MIN_N, block, grid, self.stats, best, max(1, self.elite))

  def verify(self) -> None:
    """Raise if a dependency wait of a fused generation ran out (its results are then invalid)."""
    if int(self.counters[2].item()) != 0:
      raise RuntimeError("tspmi_batch_fused_generation: a dependency wait timed out; the batch state is invalid")

  def step(self, d: BatchDraws, g: int, best=None) -> None:
    a, b = self.cur, 1 - self.cur
    s = self.shapes
    if s.get("fused"):
      self._fused(a, b, d, g, best, s["fused_block"], s["fused_grid"])
    else:
      self._staged(a, b, d, g, best, s)
    self.cur = b

  def _staged(self, a, b, d: BatchDraws, g: int, best, s: dict) -> None:
    self._generation(a, d, g, s["gen_block"], s["gen_grid"])
    _native.batch_call("tspmi_batch_survival", self.desc, self.B, self.keys[a], self.order)
    _native.batch_call("tspmi_batch_gather", self.desc, *self.w_row, self.pop[a], self.keys[a], self.off, self.order,
                       self.pop[b], self.keys[b])
    if self.elite > 0:
      self._elite_ls(self.pop[b], s["ls_block"], s["ls_grid"])
      _native.batch_call("tspmi_batch_cost", self.desc, *self.w_elite, self.pop[b], self.D, self.xy, self.ixy,
                         self.keys[b])
    if best is not None:
      _native.batch_call("tspmi_batch_best", self.desc, self.B, self.keys[b], max(1, self.elite), best)

  def run(self, d: BatchDraws, generations: int, best=None, start: int = 0) -> None:
    for t in range(generations):
      self.step(d, start + t, None if best is None else best[t])

  def capture(self, d: BatchDraws, generations: int, best=None, start: int = 0):
    """One HIP graph for ``generations`` (even) generations; a replay continues from the current state."""
    _native.require_device()
    if generations % 2:
      raise ValueError("capture needs an even number of generations")
    cur0 = self.cur
    graph = capture_graph(lambda: self.run(d, generations, best, start))
    assert self.cur == cur0
    return graph

  # ---- autotuner --------------------------------------------------------------------------------
  def autotune(self, d: BatchDraws, g: int = 0, repeats: int = 2, cache_path: str | None = None) -> dict:
    """Time the launch shapes of the two heavy stages on the live batch and keep the fastest.

    The generation stage is a pure function of the current population and draws (it only writes
    the offspring buffers), and the elite 2-opt is timed on a scratch copy, so tuning leaves the
    evolutionary state unchanged.  Results are cached per batch signature (n multiset, P, elite,
    k) in ``cache_path`` when given.
    """
    torch = _native.require_device()
    sig = f"{hash(tuple(sorted(self.ns.tolist())))}-{self.P}-{self.elite}-{self.kc}-{self.ls_iters}"
    if cache_path and os.path.exists(cache_path):
      with open(cache_path, encoding="utf-8") as fh:
        cached = json.load(fh).get(sig)
      if cached and "children_per_task" in cached:
        self.shapes = dict(cached)
        self.set_children_per_task(self.shapes["children_per_task"])
        return self.shapes
    cus = torch.cuda.get_device_properties(0).multi_processor_count

    def timed(fn) -> float:
      fn()
      torch.cuda.synchronize()
      best = float("inf")
      for _ in range(repeats):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
      return best

    a = self.cur
    b = 1 - a
    # Whole generations write only the other population buffer, the offspring buffers and the
    # offspring half of the current keys, none of which is state, so both forms can be timed on
    # the live batch without advancing it.
    fused_times = {}
    for cpt in (4, 8):
      self.set_children_per_task(cpt)
      for blk in (256, 512):
        fused_times[(blk, cpt)] = timed(lambda blk=blk: self._fused(a, b, d, g, None, blk, 0))
    self.verify()
    gen_opts = [(blk, grd) for blk in (128, 256, 512) for grd in (0, 2 * cus, 4 * cus)]
    gen_times = {o: timed(lambda o=o: self._generation(a, d, g, o[0], o[1])) for o in gen_opts}
    scratch = self.pop[a].clone()
    ls_opts = [(blk, grd) for blk in (256, 512, 1024) for grd in (0, cus, 2 * cus)]
    ls_times = {}
    for o in ls_opts:
      def run_ls(o=o):
        scratch.copy_(self.pop[a])
        self._elite_ls(scratch, o[0], o[1])
      ls_times[o] = timed(run_ls)
    gb = min(gen_times, key=gen_times.get)
    lb = min(ls_times, key=ls_times.get)
    fb = min(fused_times, key=fused_times.get)
    self.set_children_per_task(fb[1])
    shapes = {"fused": False, "fused_block": fb[0], "fused_grid": 0, "children_per_task": fb[1], "gen_block": gb[0],
              "gen_grid": gb[1], "ls_block": lb[0], "ls_grid": lb[1]}
    staged_ms = timed(lambda: self._staged(a, b, d, g, None, shapes))
    shapes["fused"] = fused_times[fb] < staged_ms
    self.shapes = shapes
    self.tuning = {"generation_ms": {f"{k[0]}x{k[1]}": v for k, v in gen_times.items()},
                   "elite_2opt_ms": {f"{k[0]}x{k[1]}": v for k, v in ls_times.items()},
                   "fused_generation_ms": {f"{k[0]}x{k[1]}": v for k, v in fused_times.items()},
                   "staged_generation_ms": staged_ms}
    if cache_path:
      data = {}
      if os.path.exists(cache_path):
        with open(cache_path, encoding="utf-8") as fh:
          data = json.load(fh)
      data[sig] = self.shapes
      with open(cache_path, "w", encoding="utf-8") as fh:
        json.dump(data, fh, indent=1)
    return self.shapes


def fused_task_list(ns, P: int, elite: int, children_per_task: int) -> np.ndarray:
  """Task words (kind << 30 | instance << 12 | index) of one fused generation, in dependency order:
  every child task, then every survive task, then the elite rows; instances longest first within
  each kind (tspmi_batch_fused_generation)."""
  ns = np.asarray(ns)
  if ns.size >= (1 << 18) or P > 4095:
    raise ValueError("the fused generation supports fewer than 2^18 instances and P < 4096")
  order = np.argsort(-ns, kind="stable").astype(np.uint32)
  e = min(int(elite), P)

  def kind(k, per):
    inst = np.repeat(order, per)
    idx = np.tile(np.arange(per, dtype=np.uint32), order.size)
    return (np.uint32(k) << np.uint32(30)) | (inst << np.uint32(12)) | idx

  parts = [kind(TASK_CHILD, -(-P // children_per_task)), kind(TASK_SURVIVE, 1)]
  if e:
    parts.append(kind(TASK_ELITE, e))
  return np.concatenate(parts).astype(np.uint32)


def c4_instances(count: int = 256, seed: int = 7, lo: int = 50, hi: int = 2000):
  """Config-4 instances: n_i uniform-integer in [lo, hi], coordinates uniform integers in [0, 1e6]^2."""
  rng = np.random.default_rng(seed)
  ns = rng.integers(lo, hi + 1, count)
  return [rng.integers(0, 1_000_001, (int(n), 2)).astype(np.float64) for n in ns]


def timed_run(ga: BatchedMemeticGA, generations: int):
  """Capture ``generations`` and replay once; returns (seconds, graph)."""
  torch = _native.require_device()
  d = ga.draw(generations)
  graph = ga.capture(d, generations)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  graph.replay()
  torch.cuda.synchronize()
  return time.perf_counter() - t0, graph
