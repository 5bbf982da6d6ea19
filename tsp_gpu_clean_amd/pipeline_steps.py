"""Extension pipeline steps (reference src/algorithms/pipeline_steps.py:36-285, registry :439-466).

``EliteOnlyLocalSearchStep`` picks the ``elite_count`` lowest-fitness rows with the device
stable rank (the reference's ``argsort(fitness)[:k]``, :163; on a fitness tie at the k-th
boundary the reference's default-kind NumPy sort is implementation-defined, the stable order is
one of its outcomes), refines them, re-evaluates and scatters them back in place.
"""

from __future__ import annotations

from typing import Any, ClassVar, Literal

import numpy as np

from . import _native
from .ga import stable_order
from .protocols import RouteAssessment, assess_with_evaluator
from .route_helpers import is_gpu_array, to_device_float64

StageAnchor = Literal["pre_selection", "pre_evaluation", "post_survival"]


def _warmup_owned(strategies: list, context) -> None:
  for s in strategies:
    hook = getattr(s, "warmup", None)
    if callable(hook):
      hook(context)


class _BaseLocalSearchPipelineStep:
  stage_anchor: ClassVar[str] = "post_survival"

  def __init__(self, local_search, evaluation) -> None:
    if local_search is None:
      raise ValueError(f"{self.__class__.__name__} requires a local_search strategy")
    if evaluation is None:
      raise ValueError(f"{self.__class__.__name__} requires an evaluation strategy")
    self.local_search = local_search
    self.evaluation = evaluation

  def strategies(self) -> list:
    return [self.local_search, self.evaluation]

  def public_params(self) -> dict[str, Any]:
    return {}

  def warmup(self, context) -> None:
    _warmup_owned(self.strategies(), context)


class PostInitLocalSearchStep(_BaseLocalSearchPipelineStep):
  """Refine the whole population once, at generation 1 (pipeline_steps.py:85-125)."""

  stage_anchor: ClassVar[str] = "pre_selection"

  def run(self, state, timer) -> None:
    if state.generation != 1:
      return
    state.population = timer.measure("local_search", self.local_search.improve_batch, state.population, state.context)
    keep_gpu = bool(is_gpu_array(state.population) and getattr(self.evaluation, "return_gpu", False))
    a = timer.measure("evaluation", assess_with_evaluator, self.evaluation, state.population, state.context,
                      return_gpu=keep_gpu)
    state.population_assessment = a
    state.fitness = a.objectives


class EliteOnlyLocalSearchStep(_BaseLocalSearchPipelineStep):
  """Refine only the elite_count best rows after survival (pipeline_steps.py:128-195)."""

  stage_anchor: ClassVar[str] = "post_survival"

  def __init__(self, local_search, evaluation, elite_count: int = 10) -> None:
    super().__init__(local_search, evaluation)
    if elite_count <= 0:
      raise ValueError(f"elite_count must be > 0, got {elite_count}")
    self.elite_count = elite_count

  def public_params(self) -> dict[str, Any]:
    return {"elite_count": self.elite_count}

  def run(self, state, timer) -> None:
    torch = _native.require_device()
    pop_size = int(state.population.shape[0])
    k = min(self.elite_count, pop_size)
    if k <= 0:
      return
    fit_dev, _ = to_device_float64(state.fitness)
    idx = stable_order(fit_dev, k)
    pop_gpu = is_gpu_array(state.population)
    if pop_gpu:
      elites = torch.empty((k, state.population.shape[1]), dtype=torch.int32, device="cuda")
      _native.gather_rows(state.population, None, None, None, idx, elites, None)
    else:
      idx_host = idx.cpu().numpy()
      elites = state.population[idx_host]
    improved = timer.measure("local_search", self.local_search.improve_batch, elites, state.context)
    keep_gpu = bool(is_gpu_array(improved) and getattr(self.evaluation, "return_gpu", False))
    ea = timer.measure("evaluation", assess_with_evaluator, self.evaluation, improved, state.context,
                       return_gpu=keep_gpu)
    if pop_gpu and is_gpu_array(improved):
      new_fit, _ = to_device_float64(ea.objectives)
      if is_gpu_array(state.fitness) and state.fitness.dtype == torch.float64 and state.fitness.is_contiguous():
        _native.scatter_rows(improved.contiguous(), new_fit, idx, state.population, state.fitness)
      else:
        _native.scatter_rows(improved.contiguous(), None, idx, state.population, None)
        fit = state.fitness if not is_gpu_array(state.fitness) else state.fitness.cpu().numpy()
        fit = np.asarray(fit, dtype=np.float64).copy()
        fit[idx.cpu().numpy()] = new_fit.cpu().numpy()
        state.fitness = fit
    else:
      idx_host = idx.cpu().numpy()
      imp = improved.cpu().numpy() if is_gpu_array(improved) else np.asarray(improved)
      obj = ea.objectives.cpu().numpy() if is_gpu_array(ea.objectives) else np.asarray(ea.objectives)
      if pop_gpu:
        d_imp = torch.from_numpy(np.ascontiguousarray(imp, dtype=np.int32)).cuda()
        _native.scatter_rows(d_imp, None, idx, state.population, None)
      else:
        state.population[idx_host] = imp
      fit = state.fitness.cpu().numpy() if is_gpu_array(state.fitness) else np.asarray(state.fitness, np.float64)
      fit = fit.copy()
      fit[idx_host] = obj
      state.fitness = fit
    pa = state.population_assessment
    if pa is not None and pa.objectives is not state.fitness:
      state.population_assessment = RouteAssessment(state.fitness, state.fitness, pa.is_feasible, pa.violations,
                                                     pa.normalized_violations, pa.raw_violation_components,
                                                     pa.num_vehicles)


class PeriodicLocalSearchStep:
  """Local search every ``interval`` generations (pipeline_steps.py:198-285)."""

  def __init__(self, local_search, interval: int = 10, target: str = "population", re_evaluate=None) -> None:
    if local_search is None:
      raise ValueError("PeriodicLocalSearchStep requires a local_search strategy")
    if interval <= 0:
      raise ValueError(f"interval must be > 0, got {interval}")
    if target not in ("offspring", "population"):
      raise ValueError(f"target must be 'offspring' or 'population', got {target!r}")
    if target == "population" and re_evaluate is None:
      raise ValueError("re_evaluate EvaluationStrategy is required when target='population'")
    self.local_search = local_search
    self.interval = interval
    self.target = target
    self.re_evaluate = re_evaluate

  @property
  def stage_anchor(self) -> str:
    return "pre_evaluation" if self.target == "offspring" else "post_survival"

  def strategies(self) -> list:
    return [self.local_search] + ([self.re_evaluate] if self.re_evaluate is not None else [])

  def public_params(self) -> dict[str, Any]:
    return {"interval": self.interval, "target": self.target}

  def warmup(self, context) -> None:
    _warmup_owned(self.strategies(), context)

  def run(self, state, timer) -> None:
    if state.generation % self.interval != 0:
      return
    if self.target == "offspring":
      if state.offspring is None:
        raise RuntimeError("PeriodicLocalSearchStep(target='offspring') called before offspring exists.")
      state.offspring = timer.measure("local_search", self.local_search.improve_batch, state.offspring, state.context)
      return
    state.population = timer.measure("local_search", self.local_search.improve_batch, state.population, state.context)
    if self.re_evaluate is not None:
      keep_gpu = bool(is_gpu_array(state.population) and getattr(self.re_evaluate, "return_gpu", False))
      a = timer.measure("evaluation", assess_with_evaluator, self.re_evaluate, state.population, state.context,
                        return_gpu=keep_gpu)
      state.population_assessment = a
      state.fitness = a.objectives


PIPELINE_STEP_REGISTRY: dict[str, type] = {
  "EliteOnlyLocalSearchStep": EliteOnlyLocalSearchStep,
  "PeriodicLocalSearchStep": PeriodicLocalSearchStep,
  "PostInitLocalSearchStep": PostInitLocalSearchStep,
}


def register_pipeline_step(name: str, step_cls: type, *, overwrite: bool = False) -> None:
  """pipeline_steps.py:446-466."""
  if not isinstance(name, str) or not name.strip():
    raise ValueError("Pipeline step name must be a non-empty string.")
  if not isinstance(step_cls, type):
    raise TypeError(f"Expected a pipeline step class (type), got {type(step_cls).__name__}.")
  if not callable(getattr(step_cls, "run", None)):
    raise TypeError(f"Pipeline step class '{step_cls.__name__}' must define a callable 'run(state, timer)' method.")
  if name in PIPELINE_STEP_REGISTRY and not overwrite:
    raise ValueError(f"Pipeline step '{name}' is already registered. Pass overwrite=True to replace it.")
  PIPELINE_STEP_REGISTRY[name] = step_cls
