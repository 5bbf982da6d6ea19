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
  4. elite 2-opt       : rank_select(k) -> gather -> tspmi_two_opt -> tspmi_tour_cost -> scatter
  5. rank_select(1)    : index of the generation's best row, kept on device for the history
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native

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


def draw_generations(rng: np.random.Generator, generations: int, P: int, n: int, k: int) -> GenerationDraws:
  """Draw G generations in the reference order from ``rng`` (advances it exactly as G steps would)."""
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
  return GenerationDraws(cand, lo, hi, probs, p1, p2)


def upload(d: GenerationDraws) -> GenerationDraws:
  import torch

  cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
  return GenerationDraws(cu(d.tour_cand), cu(d.cut_lo), cu(d.cut_hi), cu(d.probs), cu(d.pos1), cu(d.pos2))


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
    P, n = self.P, self.n
    dev = "cuda"
    self.pop = [torch.empty((P, n), dtype=torch.int32, device=dev) for _ in range(2)]
    self.keys = [torch.empty(2 * P, dtype=torch.float64, device=dev) for _ in range(2)]
    self.off = torch.empty((P, n), dtype=torch.int32, device=dev)
    self.order = torch.empty(P, dtype=torch.int32, device=dev)
    if self.elite:
      self.eidx = torch.empty(self.elite, dtype=torch.int32, device=dev)
      self.elites = torch.empty((self.elite, n), dtype=torch.int32, device=dev)
      self.efit = torch.empty(self.elite, dtype=torch.float64, device=dev)
    self.stats = torch.zeros(3, dtype=torch.int64, device=dev)
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

  def step(self, draws: GenerationDraws, g: int, best_out=None) -> None:
    """One generation using draw slot ``g``; ``best_out`` (device int32 (1,)) receives the best row."""
    P = self.P
    a, b = self.cur, 1 - self.cur
    ka, kb = self.keys[a], self.keys[b]
    _native.generation(self.pop[a], ka[:P], self.D, draws.tour_cand[g], draws.cut_lo[g], draws.cut_hi[g],
                       draws.probs[g], draws.pos1[g], draws.pos2[g], self.rate, self.off, ka[P:])
    _native.rank_select(ka, P, self.order)
    _native.gather_rows(self.pop[a], ka[:P], self.off, ka[P:], self.order, self.pop[b], kb[:P])
    if self.elite:
      _native.rank_select(kb[:P], self.elite, self.eidx)
      _native.gather_rows(self.pop[b], None, None, None, self.eidx, self.elites, None)
      if self.table is not None:
        _native.two_opt_packed(self.elites, self.D, self.table, self.cand_k, self.ls_iters, self.stats)
      else:
        _native.two_opt(self.elites, self.D, self.cand, TSPMI_MODE_BEST, self.ls_iters, self.stats)
      _native.tour_cost(self.elites, self.D, self.efit)
      _native.scatter_rows(self.elites, self.efit, self.eidx, self.pop[b], kb[:P])
    if best_out is not None:
      _native.rank_select(kb[:P], 1, best_out)
    self.cur = b

  def run(self, draws: GenerationDraws, generations: int, best_idx=None) -> None:
    """Run ``generations`` steps; ``best_idx`` (device int32 (G,)) collects each generation's best row."""
    for g in range(generations):
      self.step(draws, g, None if best_idx is None else best_idx[g : g + 1])
