"""Strategy protocols and the assessment payload (reference src/protocols/algorithm_strategies.py:62-210,
src/protocols/vrp_strategies.py:31-216).

The six PEP 544 protocols keep the reference's method names and argument order so a strategy
written against the reference registers here unchanged.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Protocol, runtime_checkable

import numpy as np

from .route_helpers import ensure_routes_batch, is_gpu_array


@runtime_checkable
class InitializationStrategy(Protocol):
  def initialize(self, customers, population_size: int, context) -> Any: ...


@runtime_checkable
class SelectionStrategy(Protocol):
  def select(self, population, fitness, context=None) -> Any: ...

  def select_batch(self, routes, fitness, num_parents: int | None = None, context=None) -> Any: ...


@runtime_checkable
class CrossoverStrategy(Protocol):
  def crossover(self, parent1, parent2, context=None) -> Any: ...

  def crossover_batch(self, routes, context=None) -> Any: ...


@runtime_checkable
class MutationStrategy(Protocol):
  def mutate(self, individual, mutation_rate: float, context=None) -> Any: ...

  def mutate_batch(self, routes, mutation_rate: float, context=None) -> Any: ...


@runtime_checkable
class LocalSearchStrategy(Protocol):
  def improve(self, route, context) -> Any: ...

  def improve_batch(self, routes, context) -> Any: ...


@runtime_checkable
class EvaluationStrategy(Protocol):
  def evaluate(self, route, context) -> float: ...

  def evaluate_batch(self, routes, context) -> Any: ...

  def canonicalize_route(self, route, context) -> Any: ...

  def assess_batch(self, routes, context, return_gpu: bool = False) -> RouteAssessment: ...


@dataclass(frozen=True)
class RouteAssessment:
  """Batch evaluation payload (vrp_strategies.py:31-54); TSP rows are feasible, zero violation."""

  objectives: Any
  travel_costs: Any
  is_feasible: Any
  violations: Any
  normalized_violations: Any
  raw_violation_components: Any | None = None
  num_vehicles: Any | None = None


def tour_assessment(costs) -> RouteAssessment:
  """Feasible assessment of Hamiltonian tours (tour_cost.py:68-82)."""
  if is_gpu_array(costs):
    import torch

    n = costs.shape[0]
    zeros = torch.zeros(n, dtype=torch.float64, device=costs.device)
    return RouteAssessment(costs, costs, torch.ones(n, dtype=torch.bool, device=costs.device), zeros,
                           zeros.clone(), torch.zeros((n, 3), dtype=torch.float64, device=costs.device),
                           torch.ones(n, dtype=torch.int32, device=costs.device))
  c = np.asarray(costs, dtype=np.float64)
  n = c.shape[0]
  zeros = np.zeros(n, dtype=np.float64)
  return RouteAssessment(c, c, np.ones(n, dtype=bool), zeros, zeros.copy(), np.zeros((n, 3)), np.ones(n, np.int32))


def assess_from_evaluate_batch(evaluator, routes, context, *, return_gpu: bool = False) -> RouteAssessment:
  """Wrap a plain evaluate_batch result (vrp_strategies.py:163-196)."""
  try:
    objectives = evaluator.evaluate_batch(routes, context, return_gpu=return_gpu)
  except TypeError:
    objectives = evaluator.evaluate_batch(routes, context)
  if return_gpu and not is_gpu_array(objectives):
    import torch

    objectives = torch.as_tensor(np.asarray(objectives, dtype=np.float64), device="cuda")
  elif not return_gpu and is_gpu_array(objectives):
    objectives = objectives.detach().cpu().numpy().astype(np.float64, copy=False)
  return tour_assessment(objectives)


def assess_with_evaluator(evaluator, routes, context, *, return_gpu: bool = False) -> RouteAssessment:
  """vrp_strategies.py:199-216."""
  batch, _ = ensure_routes_batch(routes)
  hook = getattr(evaluator, "assess_batch", None)
  if callable(hook):
    try:
      result = hook(batch, context, return_gpu=return_gpu)
    except TypeError:
      result = hook(batch, context)
    if isinstance(result, RouteAssessment):
      return result
  return assess_from_evaluate_batch(evaluator, batch, context, return_gpu=return_gpu)


def canonicalize_default_route(route) -> np.ndarray:
  """Host int32 copy (vrp_strategies.py:144-148)."""
  if is_gpu_array(route):
    return route.detach().cpu().numpy().astype(np.int32).copy()
  return np.asarray(route, dtype=np.int32).copy()


def canonicalize_with_evaluator(evaluator, route, context) -> np.ndarray:
  """vrp_strategies.py:151-160."""
  hook = getattr(evaluator, "canonicalize_route", None)
  if callable(hook):
    return np.asarray(hook(route, context), dtype=np.int32)
  return canonicalize_default_route(route)
