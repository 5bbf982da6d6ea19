"""GeneticAlgorithm with the reference's step pipeline, running its population on the MI355X.

Mirrors src/algorithms/metaheuristics/genetic_algorithm.py of the reference:
``GenerationState`` (:69-109), the step classes (:120-296), ``StageSpec``/``STAGE_SPECS``
(:304-368) and ``GeneticAlgorithm`` (:381-965) with the same constructor, pipeline composition,
incumbent rule, early stopping and stats dictionary.  Survival and elite choice run as HIP
kernels (``tspmi_rank_select`` + ``tspmi_gather_rows``): a stable ascending rank of the
concatenated fitness, which is what ``lexsort((obj, norm_viol, ~feas))`` (:682-689) reduces to
for feasible TSP tours.
"""

from __future__ import annotations

import math
from collections.abc import Callable, Mapping, Sequence
from dataclasses import dataclass
from typing import Any, Literal, Protocol

import numpy as np

from . import _native
from .protocols import RouteAssessment, assess_with_evaluator, canonicalize_with_evaluator, tour_assessment
from .route_helpers import is_gpu_array, to_device_float64, to_device_int32
from .timing import PhaseTimer


@dataclass
class GenerationState:
  population: Any
  fitness: Any
  context: Any
  offspring: Any = None
  offspring_fitness: Any = None
  generation: int = 0
  population_assessment: RouteAssessment | None = None
  offspring_assessment: RouteAssessment | None = None


class GenerationStep(Protocol):
  def run(self, state: GenerationState, timer: PhaseTimer) -> None: ...


class SelectionStep:
  def __init__(self, strategy, population_size: int) -> None:
    self.strategy = strategy
    self.population_size = population_size

  def run(self, state: GenerationState, timer: PhaseTimer) -> None:
    state.offspring = timer.measure("selection", self.strategy.select_batch, state.population, state.fitness,
                                    self.population_size, state.context)


class CrossoverStep:
  def __init__(self, strategy) -> None:
    self.strategy = strategy

  def run(self, state: GenerationState, timer: PhaseTimer) -> None:
    if state.offspring is None:
      raise RuntimeError("CrossoverStep called before state.offspring was populated.")
    state.offspring = timer.measure("crossover", self.strategy.crossover_batch, state.offspring, state.context)


class MutationStep:
  def __init__(self, strategy, mutation_rate: float) -> None:
    self.strategy = strategy
    self.mutation_rate = mutation_rate

  def run(self, state: GenerationState, timer: PhaseTimer) -> None:
    if state.offspring is None:
      raise RuntimeError("MutationStep called before state.offspring was populated.")
    state.offspring = timer.measure("mutation", self.strategy.mutate_batch, state.offspring, self.mutation_rate,
                                    state.context)


class EvaluationStep:
  def __init__(self, strategy, target: Literal["offspring", "population"] = "offspring") -> None:
    self.strategy = strategy
    self.target = target

  def run(self, state: GenerationState, timer: PhaseTimer) -> None:
    routes = state.offspring if self.target == "offspring" else state.population
    if routes is None:
      raise RuntimeError(f"EvaluationStep target={self.target!r} called with no routes.")
    keep_gpu = bool(is_gpu_array(routes) and getattr(self.strategy, "return_gpu", False))
    assessment = timer.measure("evaluation", assess_with_evaluator, self.strategy, routes, state.context,
                               return_gpu=keep_gpu)
    if self.target == "offspring":
      state.offspring_assessment = assessment
      state.offspring_fitness = assessment.objectives
    else:
      state.population_assessment = assessment
      state.fitness = assessment.objectives


class SurvivalStep:
  def __init__(self, survivor_selector) -> None:
    self.survivor_selector = survivor_selector

  def run(self, state: GenerationState, timer: PhaseTimer) -> None:
    if state.offspring is None or state.offspring_fitness is None:
      raise RuntimeError("SurvivalStep called before offspring and offspring_fitness were available.")
    owner = getattr(self.survivor_selector, "__self__", None)
    assessed = getattr(owner, "_select_survivors_assessed", None)
    if (callable(assessed) and state.population_assessment is not None and state.offspring_assessment is not None
        and getattr(self.survivor_selector, "__name__", "") == "_select_survivors"):
      state.population, state.fitness, state.population_assessment = timer.measure(
        "survival", assessed, state.population, state.population_assessment, state.offspring,
        state.offspring_assessment)
      return
    state.population, state.fitness = timer.measure("survival", self.survivor_selector, state.population,
                                                    state.fitness, state.offspring, state.offspring_fitness)
    state.population_assessment = None


class LocalSearchStep:
  def __init__(self, strategy, target: Literal["offspring", "population"] = "offspring", re_evaluate=None) -> None:
    self.strategy = strategy
    self.target = target
    self.re_evaluate = re_evaluate

  def run(self, state: GenerationState, timer: PhaseTimer) -> None:
    if self.target == "offspring":
      if state.offspring is None:
        raise RuntimeError("LocalSearchStep target='offspring' called with state.offspring is None.")
      state.offspring = timer.measure("local_search", self.strategy.improve_batch, state.offspring, state.context)
      return
    state.population = timer.measure("local_search", self.strategy.improve_batch, state.population, state.context)
    if self.re_evaluate is not None:
      keep_gpu = bool(is_gpu_array(state.population) and getattr(self.re_evaluate, "return_gpu", False))
      a = timer.measure("evaluation", assess_with_evaluator, self.re_evaluate, state.population, state.context,
                        return_gpu=keep_gpu)
      state.population_assessment = a
      state.fitness = a.objectives


@dataclass(frozen=True)
class StageSpec:
  pipeline_steps: Callable[[GeneticAlgorithm], list[GenerationStep]]
  post_loop_hook: bool = False
  requires_local_search: bool = True


def _steps_standard(ga: GeneticAlgorithm) -> list:
  return [SelectionStep(ga.selection, ga.population_size), CrossoverStep(ga.crossover),
          MutationStep(ga.mutation, ga.mutation_rate), EvaluationStep(ga.evaluation, "offspring"),
          SurvivalStep(ga._select_survivors)]


def _steps_post_crossover(ga: GeneticAlgorithm) -> list:
  if ga.local_search is None:
    return _steps_standard(ga)
  return [SelectionStep(ga.selection, ga.population_size), CrossoverStep(ga.crossover),
          LocalSearchStep(ga.local_search, "offspring"), MutationStep(ga.mutation, ga.mutation_rate),
          EvaluationStep(ga.evaluation, "offspring"), SurvivalStep(ga._select_survivors)]


def _steps_post_mutation(ga: GeneticAlgorithm) -> list:
  if ga.local_search is None:
    return _steps_standard(ga)
  return [SelectionStep(ga.selection, ga.population_size), CrossoverStep(ga.crossover),
          MutationStep(ga.mutation, ga.mutation_rate), LocalSearchStep(ga.local_search, "offspring"),
          EvaluationStep(ga.evaluation, "offspring"), SurvivalStep(ga._select_survivors)]


def _steps_post_evaluation(ga: GeneticAlgorithm) -> list:
  if ga.local_search is None:
    return _steps_standard(ga)
  return [SelectionStep(ga.selection, ga.population_size), CrossoverStep(ga.crossover),
          MutationStep(ga.mutation, ga.mutation_rate), EvaluationStep(ga.evaluation, "offspring"),
          SurvivalStep(ga._select_survivors),
          LocalSearchStep(ga.local_search, "population", re_evaluate=ga.evaluation)]


STAGE_SPECS: Mapping[str, StageSpec] = {
  "none": StageSpec(_steps_standard, post_loop_hook=False, requires_local_search=False),
  "post_crossover": StageSpec(_steps_post_crossover),
  "post_mutation": StageSpec(_steps_post_mutation),
  "post_evaluation": StageSpec(_steps_post_evaluation),
  "post_evolution": StageSpec(_steps_standard, post_loop_hook=True),
}

_PIPELINE_STRATEGY_ATTRS = ("strategy", "local_search", "evaluation", "re_evaluate")


# ---- device helpers used by survival, elites and the incumbent -------------------------------


def stable_order(keys_dev, keep: int):
  """Indices of the ``keep`` smallest keys in stable ascending order (device int32)."""
  torch = _native.require_device()
  order = torch.empty(keep, dtype=torch.int32, device="cuda")
  _native.rank_select(keys_dev, keep, order)
  return order


def _like(fit_dev, like):
  return fit_dev if is_gpu_array(like) else fit_dev.cpu().numpy()


class GeneticAlgorithm:
  def __init__(
    self,
    population_size: int = 256,
    mutation_rate: float = 0.02,
    seed: int | None = 42,
    *,
    initialization,
    selection,
    crossover,
    mutation,
    evaluation,
    local_search=None,
    local_search_stage: str = "post_mutation",
    enable_phase_timing: bool = True,
    pipeline: list | None = None,
    device_loop: bool = False,
    device_loop_chunk: int = 32,
  ) -> None:
    """The reference constructor (genetic_algorithm.py:381-457) plus ``device_loop``: when the
    pipeline is exactly the device memetic loop's generation (``engine.device_route_plan``),
    ``evolve`` runs its generations as HIP graphs of ``device_loop_chunk`` generations with one
    host read-back per chunk, with results identical to the step pipeline's; any other pipeline
    runs step by step.  On that route the strategies' own counters and timers see only the
    initialisation and the optional post-initialisation local search, and ``phase_timings`` holds
    the generations under ``"device_loop"``."""
    if local_search_stage not in STAGE_SPECS:
      raise ValueError(f"Invalid local_search_stage '{local_search_stage}'. Valid stages: {list(STAGE_SPECS)}")
    self.population_size = population_size
    self.mutation_rate = mutation_rate
    self.seed = seed
    self.local_search_stage = local_search_stage
    self.initialization = initialization
    self.selection = selection
    self.crossover = crossover
    self.mutation = mutation
    self.evaluation = evaluation
    self.local_search = local_search
    self.timer = PhaseTimer(enabled=enable_phase_timing)
    self.generation = 0
    self.best_cost_history: list[float] = []
    self.h2d_bytes = 0
    self.d2h_bytes = 0
    self.kernel_launches = 0
    self.device_loop = bool(device_loop)
    self.device_loop_chunk = int(device_loop_chunk)
    if self.device_loop and (self.device_loop_chunk < 2 or self.device_loop_chunk % 2):
      raise ValueError(f"device_loop_chunk must be an even number >= 2, got {device_loop_chunk}")
    self.custom_pipeline_steps = list(pipeline) if pipeline is not None else []
    if pipeline is not None:
      core = any(isinstance(s, (SelectionStep, CrossoverStep, MutationStep, SurvivalStep)) for s in pipeline)
      self.pipeline = pipeline if core else self._compose_pipeline(pipeline)
    else:
      self.pipeline = self._build_default_pipeline()

  def _compose_pipeline(self, extension_steps: list) -> list:
    """Insert steps at their ``stage_anchor`` (genetic_algorithm.py:459-482)."""
    steps = self._build_default_pipeline()
    pre_selection: list = []
    for step in extension_steps:
      anchor = getattr(step, "stage_anchor", "post_survival")
      if anchor == "pre_selection":
        pre_selection.append(step)
      elif anchor == "pre_evaluation":
        idx = next((i for i, s in enumerate(steps)
                    if isinstance(s, EvaluationStep) and getattr(s, "target", "") == "offspring"), len(steps))
        steps.insert(idx, step)
      else:
        steps.append(step)
    return [*pre_selection, *steps]

  def _build_default_pipeline(self) -> list:
    return STAGE_SPECS[self.local_search_stage].pipeline_steps(self)

  def _all_strategies(self) -> list:
    cands = [self.initialization, self.selection, self.crossover, self.mutation, self.evaluation, self.local_search]
    for step in self.pipeline:
      declared = getattr(step, "strategies", None)
      if callable(declared):
        cands.extend(declared())
      else:
        cands.extend(getattr(step, a, None) for a in _PIPELINE_STRATEGY_ATTRS)
    out, seen = [], set()
    for s in cands:
      if s is None or id(s) in seen:
        continue
      seen.add(id(s))
      out.append(s)
    return out

  def _reset_run_metrics(self) -> None:
    self.generation = 0
    self.best_cost_history = []
    self.h2d_bytes = self.d2h_bytes = self.kernel_launches = 0
    self.timer.reset()
    for s in self._all_strategies():
      if hasattr(s, "reset_metrics"):
        s.reset_metrics()

  def warmup(self, context) -> None:
    """Launch every kernel the pipeline can use once, without evolutionary side effects (:538-576)."""
    seen: set[int] = set()
    for step in self.pipeline:
      hook = getattr(step, "warmup", None)
      if not callable(hook):
        continue
      hook(context)
      owned = getattr(step, "strategies", None)
      if callable(owned):
        seen.update(id(s) for s in owned())
    for s in self._all_strategies():
      if id(s) in seen:
        continue
      seen.add(id(s))
      hook = getattr(s, "warmup", None)
      if callable(hook):
        hook(context)

  def _collect_strategy_metrics(self, context) -> None:
    del context
    self.h2d_bytes = sum(int(getattr(s, "h2d_bytes", 0)) for s in self._all_strategies())
    self.d2h_bytes = sum(int(getattr(s, "d2h_bytes", 0)) for s in self._all_strategies())
    self.kernel_launches = sum(int(getattr(s, "kernel_launches", 0)) for s in self._all_strategies())

  # ---- survival (device kernels) -------------------------------------------------------------
  def _survive(self, old_pop, old_fit, new_pop, new_fit):
    torch = _native.require_device()
    old_d, _, _ = to_device_int32(old_pop, copy=False)
    new_d, _, _ = to_device_int32(new_pop, copy=False)
    fo, _ = to_device_float64(old_fit)
    fn, _ = to_device_float64(new_fit)
    keys = torch.cat([fo, fn])
    order = stable_order(keys, self.population_size)
    pop = torch.empty((self.population_size, old_d.shape[1]), dtype=torch.int32, device="cuda")
    fit = torch.empty(self.population_size, dtype=torch.float64, device="cuda")
    _native.gather_rows(old_d, fo, new_d, fn, order, pop, fit)
    pop_out = pop if is_gpu_array(old_pop) else pop.cpu().numpy()
    return pop_out, fit

  def _select_survivors(self, old_pop, old_fitness, new_pop, new_fitness):
    """(mu + lambda) survival on plain fitness (:598-616), stable ascending."""
    pop, fit = self._survive(old_pop, old_fitness, new_pop, new_fitness)
    return pop, _like(fit, old_fitness)

  def _select_survivors_assessed(self, old_pop, old_assessment, new_pop, new_assessment):
    """Feasible-first survival (:618-706); TSP assessments are all feasible with zero violation."""
    pop, fit = self._survive(old_pop, old_assessment.objectives, new_pop, new_assessment.objectives)
    fit = _like(fit, old_assessment.objectives)
    return pop, fit, tour_assessment(fit)

  @staticmethod
  def _best_candidate_meta(fitness, assessment) -> tuple[int, float, bool]:
    """First minimum of the fitness; scalar read-back only (:708-739)."""
    del assessment
    if is_gpu_array(fitness):
      torch = _native.require_device()
      order = stable_order(fitness.to(torch.float64).contiguous(), 1)
      idx = int(order.item())
      return idx, float(fitness[idx].item()), True
    f = np.asarray(fitness, dtype=np.float64)
    idx = int(np.argmin(f))
    return idx, float(f[idx]), True

  @staticmethod
  def _extract_route_at(population, idx: int) -> np.ndarray:
    row = population[idx]
    if is_gpu_array(row):
      return row.cpu().numpy().astype(np.int32)
    return np.asarray(row, dtype=np.int32).copy()

  @staticmethod
  def _candidate_beats_incumbent(cand_cost, cand_feas, best_cost, best_feas, *, allow_equal: bool = False) -> bool:
    """Strict feasible-first dominance (:767-783)."""
    if cand_feas and not best_feas:
      return True
    if not cand_feas and best_feas:
      return False
    if allow_equal:
      return cand_cost <= best_cost
    return cand_cost < best_cost - 1e-6

  @staticmethod
  def _stop_reason(best_cost: float, best_feas: bool, optimal_cost, stale: int, patience: int) -> str | None:
    """Early-stopping rules checked after every generation (:894-910)."""
    if best_feas and optimal_cost is not None and optimal_cost > 0:
      if best_cost <= optimal_cost + 1e-6:
        return "hit_optimal"
      if (best_cost - optimal_cost) / optimal_cost * 100.0 < 1.0:
        return "within_1pct_optimal"
    if stale >= patience:
      return "no_improvements"
    return None

  def _evolve_on_device(self, plan: dict, state: GenerationState, max_generations: int, incumbent: dict,
                        optimal_cost, patience: int) -> None:
    """Generations of ``evolve`` on the device memetic loop; updates ``state`` and ``incumbent``."""
    from .engine import DeviceEvolution

    if plan["post_init"] is not None:
      state.generation = 1
      plan["post_init"].run(state, self.timer)
    pop_like, fit_like = state.population, state.fitness
    pop_d, _, _ = to_device_int32(state.population, copy=False)
    fit_d, _ = to_device_float64(state.fitness)
    runner = DeviceEvolution(state.context, plan, pop_d, fit_d, self.device_loop_chunk)

    def walk(first: int, costs: np.ndarray) -> int | None:
      last = stop = None
      for t, cost in enumerate(costs.tolist()):
        if self._candidate_beats_incumbent(cost, True, incumbent["cost"], incumbent["feas"]):
          incumbent["cost"], incumbent["feas"], incumbent["stale"] = cost, True, 0
          last = t
        else:
          incumbent["stale"] += 1
        self.best_cost_history.append(incumbent["cost"])
        reason = self._stop_reason(incumbent["cost"], incumbent["feas"], optimal_cost, incumbent["stale"], patience)
        if reason is not None:
          incumbent["reason"], stop = reason, t
          break
      if last is not None:
        incumbent["tour"] = runner.tour(last)
      return stop

    self.generation = self.timer.measure("device_loop", runner.run, max_generations, walk)
    self.device_loop_recoveries = runner.recoveries
    state.generation = self.generation
    pop = runner.population.clone()
    state.population = pop if is_gpu_array(pop_like) else pop.cpu().numpy()
    state.fitness = _like(runner.fitness.clone(), fit_like)
    state.population_assessment = tour_assessment(state.fitness)

  def evolve(self, context, customers: Sequence[int], max_generations: int, optimal_cost: float | None = None,
             patience: int | None = None):
    """Run the pipeline (genetic_algorithm.py:789-965); returns (best_tour, stats)."""
    if len(customers) == 0:
      raise ValueError("customers sequence cannot be empty")
    n = len(customers)
    if patience is None:
      patience = int(2 * math.sqrt(n))
    self._reset_run_metrics()
    population = self.timer.measure("initialization", self.initialization.initialize, customers,
                                    self.population_size, context)
    keep_gpu = bool(is_gpu_array(population) and getattr(self.evaluation, "return_gpu", False))
    pop_assessment = self.timer.measure("evaluation", assess_with_evaluator, self.evaluation, population, context,
                                        return_gpu=keep_gpu)
    fitness = pop_assessment.objectives
    best_idx, best_cost, best_feas = self.timer.measure("host_sync", self._best_candidate_meta, fitness, pop_assessment)
    initial_feasible_cost = best_cost
    best_tour = self.timer.measure("host_sync", self._extract_route_at, population, best_idx)
    initial_best = best_cost
    self.best_cost_history = [best_cost]
    stale = 0
    stopped_early, stop_reason = False, "max_generations"
    state = GenerationState(population=population, fitness=fitness, context=context,
                            population_assessment=pop_assessment)
    plan = None
    if self.device_loop and max_generations > 0:
      from .engine import device_route_plan

      plan = device_route_plan(self, context, population)
    if plan is not None:
      incumbent = {"cost": best_cost, "feas": best_feas, "tour": best_tour, "stale": 0, "reason": None}
      self._evolve_on_device(plan, state, max_generations, incumbent, optimal_cost, patience)
      best_cost, best_feas, best_tour = incumbent["cost"], incumbent["feas"], incumbent["tour"]
      if incumbent["reason"] is not None:
        stopped_early, stop_reason = True, incumbent["reason"]
    for gen in range(0 if plan is not None else max_generations):
      self.generation = gen + 1
      state.generation = self.generation
      for step in self.pipeline:
        step.run(state, self.timer)
      c_idx, c_cost, c_feas = self.timer.measure("host_sync", self._best_candidate_meta, state.fitness,
                                                 state.population_assessment)
      if self._candidate_beats_incumbent(c_cost, c_feas, best_cost, best_feas):
        best_tour = self.timer.measure("host_sync", self._extract_route_at, state.population, c_idx)
        best_cost, best_feas = c_cost, c_feas
        stale = 0
      else:
        stale += 1
      self.best_cost_history.append(best_cost)
      reason = self._stop_reason(best_cost, best_feas, optimal_cost, stale, patience)
      if reason is not None:
        stopped_early, stop_reason = True, reason
        break
    spec = STAGE_SPECS[self.local_search_stage]
    if self.local_search is not None and spec.post_loop_hook:
      state.population = self.timer.measure("local_search", self.local_search.improve_batch, state.population, context)
      keep_gpu = bool(is_gpu_array(state.population) and getattr(self.evaluation, "return_gpu", False))
      a = self.timer.measure("evaluation", assess_with_evaluator, self.evaluation, state.population, context,
                             return_gpu=keep_gpu)
      state.population_assessment = a
      state.fitness = a.objectives
      p_idx, p_cost, p_feas = self.timer.measure("host_sync", self._best_candidate_meta, state.fitness, a)
      if self._candidate_beats_incumbent(p_cost, p_feas, best_cost, best_feas, allow_equal=True):
        best_tour = self.timer.measure("host_sync", self._extract_route_at, state.population, p_idx)
        best_cost, best_feas = p_cost, p_feas
        self.best_cost_history[-1] = best_cost
    best_tour = canonicalize_with_evaluator(self.evaluation, best_tour, context)
    self._collect_strategy_metrics(context)
    self.final_population = state.population
    self.final_fitness = state.fitness
    stats = {
      "initial_fitness": initial_best,
      "initial_feasible_cost": initial_feasible_cost,
      "best_fitness": best_cost,
      "best_is_feasible": best_feas,
      "generations_completed": self.generation,
      "stopped_early": stopped_early,
      "stop_reason": stop_reason,
      "convergence_history": self.best_cost_history,
      "phase_timings": self.timer.summary(),
      "h2d_bytes": self.h2d_bytes,
      "d2h_bytes": self.d2h_bytes,
      "kernel_launches": self.kernel_launches,
    }
    return best_tour, stats
