"""Oracle: the reference GeneticAlgorithm.evolve() for the TSP memetic pipeline (test only).

Restates ``src/algorithms/metaheuristics/genetic_algorithm.py:789-965`` with the
pipeline that ``_compose_pipeline`` (:459-482) builds from ``local_search_stage="none"``
plus the extension steps ``PostInitLocalSearchStep`` (pipeline_steps.py:103-125,
anchor ``pre_selection``) and ``EliteOnlyLocalSearchStep`` (pipeline_steps.py:153-195,
anchor ``post_survival``):

    [PostInit?] -> Selection -> PMX -> Inversion -> Evaluation(offspring) -> Survival -> [EliteOnly?]

All random draws come from the single context generator in the reference order
(tournament_selection.py:113, pmx_crossover.py:155, inversion_mutation.py:84-92).
The incumbent is replaced only when strictly better by more than 1e-6 (:767-783);
patience defaults to int(2*sqrt(n)) (:815-816); history has one entry per completed
generation plus generation 0 (:855, :892).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import init as _init
from . import operators as ops
from . import two_opt as _two_opt


@dataclass
class LocalSearchSpec:
  mode: str = "best"
  max_iterations: int = 10
  candidates: np.ndarray | None = None  # resolved N_k table or None (exhaustive)


@dataclass
class GAResult:
  best_tour: np.ndarray
  stats: dict
  populations: list = field(default_factory=list)  # population after each generation (if recorded)
  fitnesses: list = field(default_factory=list)


def evolve(
  rng: np.random.Generator,
  dist: np.ndarray,
  population: np.ndarray,
  *,
  max_generations: int,
  mutation_rate: float,
  tournament_k: int,
  local_search: LocalSearchSpec | None = None,
  post_init: bool = False,
  elite_count: int | None = None,
  patience: int | None = None,
  optimal_cost: float | None = None,
  record: bool = False,
  crossover: str = "pmx",
  mutation: str = "inversion",
) -> GAResult:
  pop = np.array(population, dtype=np.int32, copy=True)
  p, n = pop.shape
  if patience is None:
    patience = int(2 * math.sqrt(n))
  fit = ops.tour_costs(pop, dist)
  best_idx = int(np.argmin(fit))
  best_cost = float(fit[best_idx])
  best_tour = pop[best_idx].copy()
  initial = best_cost
  history = [best_cost]
  stale = 0
  stopped_early, stop_reason = False, "max_generations"
  out = GAResult(best_tour, {})
  generation = 0

  def _ls(routes):
    ls = local_search
    return _two_opt.improve_batch(routes, dist, ls.candidates, ls.mode, ls.max_iterations)

  for gen in range(max_generations):
    generation = gen + 1
    if post_init and local_search is not None and generation == 1:
      pop = _ls(pop)
      fit = ops.tour_costs(pop, dist)
    offspring = ops.tournament_select(rng, pop, fit, tournament_k, p)
    offspring = (ops.ox_batch if crossover == "ox" else ops.pmx_batch)(rng, offspring)
    offspring = (ops.swap_batch if mutation == "swap" else ops.inversion_batch)(rng, offspring, mutation_rate)
    off_fit = ops.tour_costs(offspring, dist)
    pop, fit, _ = ops.survivors(pop, fit, offspring, off_fit, p)
    fit = fit.copy()
    if elite_count is not None and local_search is not None:
      k = min(elite_count, p)
      idx = ops.elite_indices(fit, k)
      improved = _ls(pop[idx])
      pop[idx] = improved
      fit[idx] = ops.tour_costs(improved, dist)
    if record:
      out.populations.append(pop.copy())
      out.fitnesses.append(fit.copy())
    c_idx = int(np.argmin(fit))
    c_cost = float(fit[c_idx])
    if c_cost < best_cost - 1e-6:
      best_tour = pop[c_idx].copy()
      best_cost = c_cost
      stale = 0
    else:
      stale += 1
    history.append(best_cost)
    if optimal_cost is not None and optimal_cost > 0:
      if best_cost <= optimal_cost + 1e-6:
        stopped_early, stop_reason = True, "hit_optimal"
        break
      if (best_cost - optimal_cost) / optimal_cost * 100.0 < 1.0:
        stopped_early, stop_reason = True, "within_1pct_optimal"
        break
    if stale >= patience:
      stopped_early, stop_reason = True, "no_improvements"
      break

  out.best_tour = best_tour
  out.stats = {
    "initial_fitness": initial,
    "best_fitness": best_cost,
    "generations_completed": generation,
    "stopped_early": stopped_early,
    "stop_reason": stop_reason,
    "convergence_history": history,
  }
  return out


def init_population(rng, dist, n, pop_size, kind="random", k=5):
  """Row seeds then rows, random_init.py:42-55 / nearest_neighbor_init.py:124-132."""
  seeds = _init.draw_row_seeds(rng, pop_size)
  customers = np.arange(n, dtype=np.int32)
  if kind == "random":
    return _init.random_init(customers, pop_size, seeds)
  from .distances import compute_candidate_list

  return _init.nn_init(customers, pop_size, compute_candidate_list(dist, k), seeds)
