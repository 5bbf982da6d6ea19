"""Device-resident memetic generation loop: the reference pipeline
``Selection -> PMX -> Inversion -> Evaluation -> Survival -> EliteOnly 2-opt`` with no host
synchronisation inside the loop.

Every random number is drawn on the host beforehand, in the order the reference strategies
consume ``context.rng`` each generation (tournament_selection.py:113, pmx_crossover.py:155,
inversion_mutation.py:84-92); none of those draws depends on the population, so drawing all
generations up front leaves the stream, and therefore every generation, bit-identical to the
step-by-step ``GeneticAlgorithm`` pipeline.  Per generation the device runs:

  1. tspmi_generation  : fused tournament -> PMX -> inversion -> cost into the offspring buffer
  2. tspmi_rank_select : stable (mu + lambda) order of concat(fit, off_fit)
  3. tspmi_gather_rows : survivors into the other population buffer (ping-pong)
  4. elite 2-opt       : tspmi_two_opt_elite (search and re-score in one launch, one cluster of
                         workgroups per tour) on the sorted survivors' prefix
  5. best row          : index of the generation's best row, kept on device for the history (taken
                         by the elite launch when it runs as workgroup clusters)
Whole runs of generations can be captured into one HIP graph (``capture``).
"""

from __future__ import annotations

from dataclasses import dataclass, fields

import numpy as np

from . import _native
from .context import coordinate_candidates_pay

TSPMI_MODE_BEST = 0


@dataclass
class GenerationDraws:
  """Host draws of G generations, stacked (device tensors after ``upload``)."""

  tour_cand: object  # (G, P, k) int32
  cut_lo: object  # (G, max(1, P//2)) int32
  cut_hi: object
  probs: object  # (G, P) float64
  pos1: object  # (G, P) int32
  pos2: object


def capture_graph(fn):
  """Capture the device work ``fn()`` enqueues into one HIP graph (replayed on the current stream).

  Captures with ``capture_begin`` on a side stream rather than the ``torch.cuda.graph`` context,
  whose entry runs a Python garbage collection and empties torch's allocator cache (tens of ms in
  a large process); ``fn`` must not allocate device memory.
  """
  torch = _native.require_device()
  graph = torch.cuda.CUDAGraph()
  side = torch.cuda.Stream()
  side.wait_stream(torch.cuda.current_stream())
  with torch.cuda.stream(side):
    graph.capture_begin()
    try:
      fn()
    finally:
      graph.capture_end()
  torch.cuda.current_stream().wait_stream(side)
  return graph


def draw_generations(rng: np.random.Generator, generations: int, P: int, n: int, k: int,
                     states: list | None = None) -> GenerationDraws:
  """Draw G generations in the reference order from ``rng`` (advances it exactly as G steps would);
  ``states`` receives the generator state after each generation's draws.

  A PCG64 generator (NumPy's default) is drawn by ``tspmi_draw_generations`` in one call, with
  the same values and end state as the per-call NumPy loop below, which serves other bit
  generators.
  """
  state = rng.bit_generator.state
  if state.get("bit_generator") == "PCG64" and generations > 0:
    new, cand, lo, hi, probs, p1, p2, words = _native.draw_generations_pcg64(state, generations, P, n, k,
                                                                             states is not None)
    rng.bit_generator.state = new
    if states is not None:
      states.extend(_native.pcg64_state(w, state) for w in words)
    return GenerationDraws(cand, lo, hi, probs, p1, p2)
  return draw_generations_numpy(rng, generations, P, n, k, states)


def draw_generations_numpy(rng: np.random.Generator, generations: int, P: int, n: int, k: int,
                           states: list | None = None) -> GenerationDraws:
  """``draw_generations`` by one NumPy call per reference draw (any bit generator)."""
  half = P // 2
  cand = np.empty((generations, P, k), np.int32)
  lo = np.zeros((generations, max(1, half)), np.int32)
  hi = np.zeros((generations, max(1, half)), np.int32)
  probs = np.empty((generations, P), np.float64)
  p1 = np.empty((generations, P), np.int32)
  p2 = np.empty((generations, P), np.int32)
  for g in range(generations):
    cand[g] = rng.integers(0, P, size=(P, k), dtype=np.int32)
    if half > 0:
      first = rng.integers(0, n, size=half, dtype=np.int32)
      sh = rng.integers(0, n - 1, size=half, dtype=np.int32)
      second = sh + (sh >= first).astype(np.int32)
      lo[g, :half] = np.minimum(first, second)
      hi[g, :half] = np.maximum(first, second)
    probs[g] = rng.random(P)
    first = rng.integers(0, n, size=P, dtype=np.int32)
    sh = rng.integers(0, n - 1, size=P, dtype=np.int32)
    p1[g] = first
    p2[g] = sh + (sh >= first).astype(np.int32)
    if states is not None:
      states.append(rng.bit_generator.state)
  return GenerationDraws(cand, lo, hi, probs, p1, p2)


def upload(d: GenerationDraws) -> GenerationDraws:
  import torch

  cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
  return GenerationDraws(cu(d.tour_cand), cu(d.cut_lo), cu(d.cut_hi), cu(d.probs), cu(d.pos1), cu(d.pos2))


def upload_into(dst: GenerationDraws, src: GenerationDraws) -> None:
  """Copy host draws of g generations into slots 0 .. g-1 of device draws of the same shape."""
  import torch

  g = int(src.tour_cand.shape[0])
  for f in fields(GenerationDraws):
    getattr(dst, f.name)[:g].copy_(torch.from_numpy(np.ascontiguousarray(getattr(src, f.name))))


class DeviceMemeticLoop:
  """Runs generations of the C1/C3 memetic pipeline entirely on the device."""

  def __init__(self, context, population_size: int, mutation_rate: float, tournament_k: int = 5,
               elite_count: int | None = 16, ls_max_iterations: int = 10, ls_candidate_k: int | None = 20) -> None:
    torch = _native.require_device()
    self.context = context
    self.P = int(population_size)
    self.n = int(context.dimension)
    self.rate = float(mutation_rate)
    self.k = int(tournament_k)
    self.elite = None if elite_count is None else min(int(elite_count), self.P)
    self.ls_iters = int(ls_max_iterations)
    self.D = context.get_gpu_distances()
    self.cand = context.get_gpu_candidate_list(ls_candidate_k) if ls_candidate_k else None
    self.cand_k = 0 if self.cand is None else int(self.cand.shape[1])
    self.table = (context.get_gpu_candidate_table(ls_candidate_k)
                  if self.cand is not None and 1 <= self.cand_k <= 32 else None)
    # Above a few thousand cities the candidate scan computes distances from the coordinates.
    self.xy = context.get_gpu_euc2d_coordinates() if coordinate_candidates_pay(self.n) else None
    if self.xy is not None and context.get_gpu_int_coordinates() is not None:
      self.xy = context.get_gpu_int_coordinates()  # integer instances: the integer distance form
    # The elite clusters compute their distances from int32 coordinates where the instance has them
    # (no matrix gathers in a pass).
    self.ixy = context.get_gpu_int_coordinates()
    P, n = self.P, self.n
    dev = "cuda"
    self.pop = [torch.empty((P, n), dtype=torch.int32, device=dev) for _ in range(2)]
    self.keys = [torch.empty(2 * P, dtype=torch.float64, device=dev) for _ in range(2)]
    self.off = torch.empty((P, n), dtype=torch.int32, device=dev)
    self.order = torch.empty(P, dtype=torch.int32, device=dev)
    self.stats = torch.zeros(3, dtype=torch.int64, device=dev)
    # Barrier words of the elite search's workgroup clusters (left zeroed by every launch).
    self.sync = torch.zeros(_native.ELITE_SYNC_WORDS * max(1, self.elite or 1), dtype=torch.int64, device=dev)
    # False: the elite search runs one workgroup per tour (same results, no cluster barrier); the
    # recovery path after a cluster could not meet.
    self.clusters = True
    self.cur = 0

  def load(self, population, fitness=None) -> None:
    """Set the current population (device int32 (P, n)); costs are evaluated when not given."""
    self.pop[self.cur].copy_(population)
    fit = self.keys[self.cur][: self.P]
    if fitness is None:
      _native.tour_cost(self.pop[self.cur], self.D, fit)
    else:
      fit.copy_(fitness)

  @property
  def population(self):
    return self.pop[self.cur]

  @property
  def fitness(self):
    return self.keys[self.cur][: self.P]

  def step(self, draws: GenerationDraws, g: int, best_out=None, best_tour=None, best_cost=None) -> None:
    """One generation using draw slot ``g``; ``best_out`` (device int32 (1,)) receives the best row,
    and ``best_tour`` (int32 (1, n)) / ``best_cost`` (float64 (1,)), when given, its tour and cost.

    Survivors are gathered in their stable rank order, so the elite set of the reference's
    ``argsort(fitness)[:elite]`` (pipeline_steps.py:163) is simply rows 0..elite-1: the elite
    2-opt refines that prefix in place and re-scores it into the same key slots.  Because local
    search never raises a cost, the generation's first minimum then lies within that prefix.
    """
    P = self.P
    a, b = self.cur, 1 - self.cur
    ka, kb = self.keys[a], self.keys[b]
    _native.generation(self.pop[a], ka[:P], self.D, draws.tour_cand[g], draws.cut_lo[g], draws.cut_hi[g],
                       draws.probs[g], draws.pos1[g], draws.pos2[g], self.rate, self.off, ka[P:])
    _native.survive(ka, P, self.pop[a], self.off, self.pop[b], kb[:P], self.order)
    best_done = False
    if self.elite:
      E = self.elite
      elites = self.pop[b][:E]
      fused = self.table is not None and self.n >= 4 and self.ls_iters != 0
      if fused and self.xy is not None:
        _native.two_opt_packed_euc2d(elites, self.xy, self.table, self.cand_k, self.ls_iters, self.stats, kb[:E])
      elif fused and self.clusters:
        _native.two_opt_elite(elites, self.D, self.table, self.cand_k, self.ls_iters, self.stats, kb[:E], self.sync,
                              best_out, ixy=self.ixy)
        best_done = True
      elif fused:
        _native.two_opt_packed(elites, self.D, self.table, self.cand_k, self.ls_iters, self.stats, kb[:E])
      else:
        if self.table is not None:
          _native.two_opt_packed(elites, self.D, self.table, self.cand_k, self.ls_iters, self.stats)
        else:
          _native.two_opt(elites, self.D, self.cand, TSPMI_MODE_BEST, self.ls_iters, self.stats)
        _native.tour_cost(elites, self.D, kb[:E])
    if best_out is not None and not best_done:
      _native.rank_select(kb[: (self.elite or 1)], 1, best_out)
    if best_tour is not None:
      _native.gather_rows(self.pop[b], kb[:P], None, None, best_out, best_tour, best_cost)
    self.cur = b

  def elite_failed(self) -> bool:
    """Whether an elite search's workgroup cluster could not meet since the last clear (that tour
    was left unrefined, so the generations from it on differ from the reference); reads the
    device, so call it after the work has been issued and outside graph capture."""
    return bool(self.elite) and _native.elite_failed(self.sync)

  def clear_elite_failure(self) -> None:
    self.sync.zero_()

  def verify(self) -> None:
    """Raise if an elite search's workgroup cluster could not meet (its result was not written)."""
    if self.elite_failed():
      raise RuntimeError("elite 2-opt: a workgroup cluster was not co-resident; results are invalid")

  def rerun_without_clusters(self, population, fitness, draws: GenerationDraws, generations: int, best_idx=None,
                             best_tours=None, best_costs=None) -> None:
    """Run ``generations`` steps again from ``population`` / ``fitness`` with the elite search on
    one workgroup per tour (no cluster barrier), after clearing the failure; same results as a run
    whose clusters all met."""
    self.clear_elite_failure()
    self.load(population, fitness)
    self.clusters = False
    try:
      self.run(draws, generations, best_idx, 0, best_tours, best_costs)
    finally:
      self.clusters = True

  def run(self, draws: GenerationDraws, generations: int, best_idx=None, start: int = 0, best_tours=None,
          best_costs=None) -> None:
    """Run ``generations`` steps from draw slot ``start``; ``best_idx`` collects each best row and
    ``best_tours`` (G, n) / ``best_costs`` (G,) its tour and cost (they need ``best_idx``)."""
    if best_tours is not None and best_idx is None:
      raise ValueError("best_tours needs best_idx")
    for t in range(generations):
      row = slice(t, t + 1)
      self.step(draws, start + t, None if best_idx is None else best_idx[row],
                None if best_tours is None else best_tours[row], None if best_costs is None else best_costs[row])

  def capture(self, draws: GenerationDraws, generations: int, best_idx=None, start: int = 0, best_tours=None,
              best_costs=None):
    """Capture ``generations`` (even) steps into one HIP graph; replaying it advances the loop.

    The graph reads the buffers of the current population slot and, the count being even, ends
    in the same slot, so every replay continues from the state the previous one left.  Each
    replay consumes draw slots start .. start + generations - 1 again: refill them (or capture a
    whole run) to keep the random stream advancing.
    """
    _native.require_device()
    if generations % 2:
      raise ValueError("capture needs an even number of generations")
    cur0 = self.cur
    graph = capture_graph(lambda: self.run(draws, generations, best_idx, start, best_tours, best_costs))
    assert self.cur == cur0
    return graph


#: Largest n the fused generation kernel holds in LDS (six int32 rows of n, 160 KiB per CU).
GENERATION_MAX_N = (160 * 1024 - 64) // 24


def device_route_plan(ga, context, population) -> dict | None:
  """Settings for ``DeviceMemeticLoop`` when ``ga``'s generation is exactly the loop's, else None.

  The loop reproduces the pipeline ``[PostInitLocalSearchStep] -> tournament selection -> PMX ->
  inversion -> offspring evaluation -> (mu + lambda) survival -> [EliteOnlyLocalSearchStep]``
  (genetic_algorithm.py:459-482 composition, pipeline_steps.py:85-195) with every strategy
  drawing from the context generator, whole-instance permutations, and an elite 2-opt that takes
  the steepest-move candidate kernel.  Any other pipeline, strategy, option or shape returns
  None and the step pipeline runs instead.
  """
  from .ga import CrossoverStep, EvaluationStep, MutationStep, SelectionStep, SurvivalStep
  from .pipeline_steps import EliteOnlyLocalSearchStep, PostInitLocalSearchStep
  from .route_helpers import resolve_route_bounds
  from .strategies.cpu_named import (CrossoverCPUPermPartiallyMapped, EvaluationCPUCost, LocalSearchCPUTourTwoOpt,
                                     MutationCPUPermInversion, SelectionCPUTournament)
  from .strategies.crossover import CrossoverGPUPermPartiallyMapped
  from .strategies.evaluation import EvaluationGPUCost
  from .strategies.local_search import _MODES, LocalSearchGPUTourTwoOpt
  from .strategies.mutation import MutationGPUPermInversion
  from .strategies.selection import SelectionGPUTournament

  evaluations = (EvaluationGPUCost, EvaluationCPUCost)
  if ga.local_search is not None and ga.local_search_stage != "none":
    return None
  if not isinstance(getattr(context, "rng", None), np.random.Generator):
    return None
  if type(ga.selection) not in (SelectionGPUTournament, SelectionCPUTournament) or int(ga.selection.k) < 1:
    return None
  if type(ga.crossover) not in (CrossoverGPUPermPartiallyMapped, CrossoverCPUPermPartiallyMapped):
    return None
  if type(ga.mutation) not in (MutationGPUPermInversion, MutationCPUPermInversion):
    return None
  if type(ga.evaluation) not in evaluations:
    return None
  P, n = int(ga.population_size), int(context.dimension)
  shape = tuple(int(x) for x in getattr(population, "shape", ()))
  if shape != (P, n) or P < 2 or P % 2 or n < 4 or n > GENERATION_MAX_N:
    return None
  if resolve_route_bounds(n, context=context) != (0, n):
    return None
  steps = list(ga.pipeline)
  post_init = steps.pop(0) if steps and type(steps[0]) is PostInitLocalSearchStep else None
  elite_step = steps.pop() if steps and type(steps[-1]) is EliteOnlyLocalSearchStep else None
  core = [SelectionStep, CrossoverStep, MutationStep, EvaluationStep, SurvivalStep]
  if [type(s) for s in steps] != core:
    return None
  sel, cx, mut, ev, surv = steps
  if (sel.strategy is not ga.selection or sel.population_size != P or cx.strategy is not ga.crossover
      or mut.strategy is not ga.mutation or mut.mutation_rate != ga.mutation_rate or ev.strategy is not ga.evaluation
      or ev.target != "offspring" or surv.survivor_selector != ga._select_survivors):
    return None
  plan = {"P": P, "rate": float(ga.mutation_rate), "k": int(ga.selection.k), "elite": None, "iters": 0,
          "cand_k": None, "post_init": post_init}
  if elite_step is not None:
    ls = elite_step.local_search
    if (type(ls) not in (LocalSearchGPUTourTwoOpt, LocalSearchCPUTourTwoOpt) or _MODES[ls.mode] != TSPMI_MODE_BEST
        or ls.candidate_scorer is not None or int(ls.max_iterations) <= 0 or not ls._uses_candidates(context, n)
        or not 1 <= int(ls.candidate_k) <= 32 or type(elite_step.evaluation) not in evaluations):
      return None
    plan.update(elite=min(int(elite_step.elite_count), P), iters=int(ls.max_iterations), cand_k=int(ls.candidate_k))
  return plan


class DeviceEvolution:
  """Runs ``GeneticAlgorithm.evolve``'s generations on ``DeviceMemeticLoop`` in chunks of
  ``chunk`` generations, each replayed from one captured HIP graph.

  A chunk's draws are made on the host, in the reference order from the context generator,
  while the previous chunk runs; the first chunk runs eagerly while the graph is captured.  After
  a chunk the best cost of each of its generations is read back and handed to the caller, which
  applies the incumbent and stopping rules
  (genetic_algorithm.py:876-910) and names the generation to stop at, if any.  A stop inside a
  chunk rewinds: the population saved at the chunk start is run eagerly to the stopping
  generation and the generator is set to its state after that generation's draws, so population,
  fitness and generator end where the step pipeline ends.  A chunk in which a cluster of the elite
  search could not meet (its bounded barrier timed out, leaving a tour unrefined) is run again
  from its start with the elite search on one workgroup per tour, which gives the same results.
  """

  def __init__(self, context, plan: dict, population, fitness, chunk: int = 32) -> None:
    torch = _native.require_device()
    if chunk < 2 or chunk % 2:
      raise ValueError(f"chunk must be an even number >= 2, got {chunk}")
    self.loop = DeviceMemeticLoop(context, plan["P"], plan["rate"], plan["k"], plan["elite"], plan["iters"],
                                  plan["cand_k"])
    self.loop.load(population, fitness)
    self.rng = context.rng
    self.chunk = int(chunk)
    P, n, C, k = self.loop.P, self.loop.n, self.chunk, plan["k"]
    half = max(1, P // 2)
    dev = "cuda"
    self.draws = GenerationDraws(torch.empty((C, P, k), dtype=torch.int32, device=dev),
                                 torch.empty((C, half), dtype=torch.int32, device=dev),
                                 torch.empty((C, half), dtype=torch.int32, device=dev),
                                 torch.empty((C, P), dtype=torch.float64, device=dev),
                                 torch.empty((C, P), dtype=torch.int32, device=dev),
                                 torch.empty((C, P), dtype=torch.int32, device=dev))
    self.best_idx = torch.empty(C, dtype=torch.int32, device=dev)
    self.best_tours = torch.empty((C, n), dtype=torch.int32, device=dev)
    self.best_costs = torch.empty(C, dtype=torch.float64, device=dev)
    self.saved_pop = torch.empty((P, n), dtype=torch.int32, device=dev)
    self.saved_fit = torch.empty(P, dtype=torch.float64, device=dev)
    self.graph = None
    self.k = k
    self.recoveries = 0  # chunks re-run after an elite cluster could not meet

  def _draw(self, generations: int):
    states: list = []
    host = draw_generations(self.rng, generations, self.loop.P, self.loop.n, self.k, states)
    return host, states

  def tour(self, t: int) -> np.ndarray:
    """Best tour of row ``t`` of the chunk last handed to the caller."""
    return self.best_tours[t].cpu().numpy()

  def run(self, max_generations: int, walk) -> int:
    """Run up to ``max_generations``; ``walk(first_generation, costs)`` gets each chunk's best
    costs (float64 per generation) and returns the row to stop at, or None.  Returns the number of
    generations completed."""
    done = 0
    pending = self._draw(min(self.chunk, max_generations)) if max_generations > 0 else None
    while pending is not None:
      host, states = pending
      g = int(host.tour_cand.shape[0])
      upload_into(self.draws, host)
      self.saved_pop.copy_(self.loop.population)
      self.saved_fit.copy_(self.loop.fitness)
      if g == self.chunk and self.graph is not None:
        self.graph.replay()
      else:  # the first chunk, and a short last one, run eagerly
        self.loop.run(self.draws, g, self.best_idx[:g], 0, self.best_tours[:g], self.best_costs[:g])
      left = max_generations - done - g
      if self.graph is None and left >= self.chunk:
        # capturing records work without running it: it overlaps the eager chunk on the device
        self.graph = self.loop.capture(self.draws, self.chunk, self.best_idx, 0, self.best_tours, self.best_costs)
      pending = self._draw(min(self.chunk, left)) if left > 0 else None  # overlaps the device work
      costs = self.best_costs[:g].cpu().numpy()
      if self.loop.elite_failed():
        # a cluster of the elite search could not meet: the chunk from that generation on is not
        # the reference's; run it again from its start with one workgroup per elite tour
        self.recoveries += 1
        self.loop.rerun_without_clusters(self.saved_pop, self.saved_fit, self.draws, g, self.best_idx[:g],
                                         self.best_tours[:g], self.best_costs[:g])
        costs = self.best_costs[:g].cpu().numpy()
      stop = walk(done, costs)
      if stop is not None:
        if stop + 1 < g:
          self.loop.load(self.saved_pop, self.saved_fit)
          self.loop.run(self.draws, stop + 1)
          if self.loop.elite_failed():
            self.recoveries += 1
            self.loop.rerun_without_clusters(self.saved_pop, self.saved_fit, self.draws, stop + 1)
        self.rng.bit_generator.state = states[stop]
        return done + stop + 1
      done += g
    return done

  @property
  def population(self):
    return self.loop.population

  @property
  def fitness(self):
    return self.loop.fitness
