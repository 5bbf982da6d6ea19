"""Strategy registry (reference src/algorithms/strategies/__init__.py:87-207).

The reference's CPU strategy names for TSP tours are registered too (``cpu_named``): they run on
the device with the CPU classes' signatures and give the same results, as the reference's own CPU
and GPU classes do for one seed.

``register_strategy(name, cls)`` validates reflectively that ``cls`` implements every callable
of at least one of the six protocols, exactly like the reference, so third-party strategies
and the HIP ones plug into the same slots.
"""

from __future__ import annotations

from ..protocols import (
  CrossoverStrategy,
  EvaluationStrategy,
  InitializationStrategy,
  LocalSearchStrategy,
  MutationStrategy,
  SelectionStrategy,
)
from .base import StrategyBase
from .cpu_named import (
  CrossoverCPUPermOrder,
  CrossoverCPUPermPartiallyMapped,
  EvaluationCPUCost,
  InitializationCPUPermRandom,
  InitializationCPUTourNearestNeighbor,
  LocalSearchCPUTourTwoOpt,
  MutationCPUPermInversion,
  MutationCPUPermSwap,
  SelectionCPUTournament,
)
from .crossover import CrossoverGPUPermOrder, CrossoverGPUPermPartiallyMapped
from .evaluation import EvaluationGPUCost
from .initialization import InitializationGPUPermRandom, InitializationGPUTourNearestNeighbor
from .local_search import LocalSearchGPUTourTwoOpt, LocalSearchNoOp
from .mutation import MutationGPUPermInversion, MutationGPUPermSwap
from .selection import SelectionGPUTournament

STRATEGY_REGISTRY: dict[str, type] = {
  "CrossoverCPUPermOrder": CrossoverCPUPermOrder,
  "CrossoverCPUPermPartiallyMapped": CrossoverCPUPermPartiallyMapped,
  "EvaluationCPUCost": EvaluationCPUCost,
  "InitializationCPUPermRandom": InitializationCPUPermRandom,
  "InitializationCPUTourNearestNeighbor": InitializationCPUTourNearestNeighbor,
  "LocalSearchCPUTourTwoOpt": LocalSearchCPUTourTwoOpt,
  "MutationCPUPermInversion": MutationCPUPermInversion,
  "MutationCPUPermSwap": MutationCPUPermSwap,
  "SelectionCPUTournament": SelectionCPUTournament,
  "CrossoverGPUPermPartiallyMapped": CrossoverGPUPermPartiallyMapped,
  "CrossoverGPUPermOrder": CrossoverGPUPermOrder,
  "EvaluationGPUCost": EvaluationGPUCost,
  "InitializationGPUPermRandom": InitializationGPUPermRandom,
  "InitializationGPUTourNearestNeighbor": InitializationGPUTourNearestNeighbor,
  "LocalSearchGPUTourTwoOpt": LocalSearchGPUTourTwoOpt,
  "LocalSearchNoOp": LocalSearchNoOp,
  "MutationGPUPermInversion": MutationGPUPermInversion,
  "MutationGPUPermSwap": MutationGPUPermSwap,
  "SelectionGPUTournament": SelectionGPUTournament,
}

_STRATEGY_PROTOCOLS: tuple[type, ...] = (
  InitializationStrategy,
  SelectionStrategy,
  CrossoverStrategy,
  MutationStrategy,
  LocalSearchStrategy,
  EvaluationStrategy,
)

_STRATEGY_PROTOCOL_METHODS: dict[str, tuple[str, ...]] = {
  proto.__name__: tuple(sorted(name for name, attr in vars(proto).items() if not name.startswith("_") and callable(attr)))
  for proto in _STRATEGY_PROTOCOLS
}


def register_strategy(name: str, strategy_cls: type, *, overwrite: bool = False) -> None:
  """Register a strategy class (reference strategies/__init__.py:175-207)."""
  if not isinstance(name, str) or not name.strip():
    raise ValueError("Strategy name must be a non-empty string.")
  if not isinstance(strategy_cls, type):
    raise TypeError(f"Expected a strategy class (type), got {type(strategy_cls).__name__}.")
  if not any(all(callable(getattr(strategy_cls, m, None)) for m in methods)
             for methods in _STRATEGY_PROTOCOL_METHODS.values()):
    expected = ", ".join(f"{p}({', '.join(m)})" for p, m in _STRATEGY_PROTOCOL_METHODS.items())
    raise TypeError(f"Strategy class '{strategy_cls.__name__}' does not implement any strategy protocol. "
                    f"Expected one of: {expected}.")
  if name in STRATEGY_REGISTRY and not overwrite:
    raise ValueError(f"Strategy '{name}' is already registered. Pass overwrite=True to replace it.")
  STRATEGY_REGISTRY[name] = strategy_cls


__all__ = [
  "STRATEGY_REGISTRY",
  "CrossoverCPUPermOrder",
  "CrossoverCPUPermPartiallyMapped",
  "EvaluationCPUCost",
  "InitializationCPUPermRandom",
  "InitializationCPUTourNearestNeighbor",
  "LocalSearchCPUTourTwoOpt",
  "MutationCPUPermInversion",
  "MutationCPUPermSwap",
  "SelectionCPUTournament",
  "CrossoverGPUPermPartiallyMapped",
  "CrossoverGPUPermOrder",
  "EvaluationGPUCost",
  "InitializationGPUPermRandom",
  "InitializationGPUTourNearestNeighbor",
  "LocalSearchGPUTourTwoOpt",
  "LocalSearchNoOp",
  "MutationGPUPermInversion",
  "MutationGPUPermSwap",
  "SelectionGPUTournament",
  "StrategyBase",
  "register_strategy",
]
