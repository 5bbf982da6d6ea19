"""The reference's CPU strategy classes for TSP tours, under their own names, on the device.

Each reference CPU class draws from ``context.rng`` exactly as its GPU twin does -- the draws are
shared helpers (``rng.integers(0, P, (N, k))`` for tournaments, one ``sample_distinct_pairs``
call per crossover batch, ``random(P)`` before the mutation positions, ``draw_row_seeds`` for
both initialisers) and both backends read the same candidate lists -- and returns its result where
the input lives.  So a configuration that names them (the reference's "CPU" and "Hybrid*" entries
of tests/fixtures/configs/algorithms.json) runs unchanged here, through the same HIP kernels as
the GPU-named classes, with the CPU classes' constructor signatures, defaults and validation.
VRP-only options (decoders, penalty policies) are rejected.
"""

from __future__ import annotations

from typing import Any

from .crossover import CrossoverGPUPermOrder, CrossoverGPUPermPartiallyMapped
from .evaluation import EvaluationGPUCost
from .initialization import InitializationGPUPermRandom, InitializationGPUTourNearestNeighbor
from .local_search import LocalSearchGPUTourTwoOpt
from .mutation import MutationGPUPermInversion, MutationGPUPermSwap
from .selection import SelectionGPUTournament

_CPU_TWO_OPT_MODES = ("best", "first", "multi_move")


class SelectionCPUTournament(SelectionGPUTournament):
  """tournament_selection.py:29-128 (default k = 3)."""

  def __init__(self, k: int = 3, seed: Any = None) -> None:
    super().__init__(k=k, seed=seed)


class CrossoverCPUPermPartiallyMapped(CrossoverGPUPermPartiallyMapped):
  """pmx_crossover.py:36-171."""


class CrossoverCPUPermOrder(CrossoverGPUPermOrder):
  """order_crossover.py:36-160."""


class MutationCPUPermInversion(MutationGPUPermInversion):
  """inversion_mutation.py:36-99."""


class MutationCPUPermSwap(MutationGPUPermSwap):
  """swap_mutation.py:36-99."""


class InitializationCPUPermRandom(InitializationGPUPermRandom):
  """random_init.py:24-55 (host-resident result)."""

  def __init__(self, seed: Any = None) -> None:
    super().__init__(seed=seed, return_gpu=False)


class InitializationCPUTourNearestNeighbor(InitializationGPUTourNearestNeighbor):
  """nearest_neighbor_init.py:102-132 (host-resident result)."""

  def __init__(self, k: int = 20, seed: Any = None) -> None:
    super().__init__(k=k, seed=seed, return_gpu=False)


class EvaluationCPUCost(EvaluationGPUCost):
  """cost_evaluation.py:46-162 for TSP tours (host-resident costs)."""

  def __init__(self, decoder: Any = None, split_mode: str = "prins", penalty_policy: Any = None,
               max_matrix_dimension: int | None = None) -> None:
    if decoder is not None or penalty_policy is not None:
      raise NotImplementedError("EvaluationCPUCost: VRP decoders and penalty policies are not implemented")
    super().__init__(return_gpu=False, max_matrix_dimension=max_matrix_dimension, split_mode=split_mode)


class LocalSearchCPUTourTwoOpt(LocalSearchGPUTourTwoOpt):
  """two_opt.py:264-628: modes best (default), first, multi_move; max_iterations >= 0."""

  def __init__(self, max_iterations: int = 10, mode: str = "best", candidate_k: int | None = None) -> None:
    if mode not in _CPU_TWO_OPT_MODES:
      raise ValueError(f"Invalid 2-opt CPU mode {mode!r}; expected one of {sorted(_CPU_TWO_OPT_MODES)}")
    if max_iterations < 0:
      raise ValueError(f"max_iterations must be >= 0, got {max_iterations}")
    super().__init__(max_iterations=max_iterations, mode=mode, candidate_k=candidate_k)


__all__ = [
  "CrossoverCPUPermOrder",
  "CrossoverCPUPermPartiallyMapped",
  "EvaluationCPUCost",
  "InitializationCPUPermRandom",
  "InitializationCPUTourNearestNeighbor",
  "LocalSearchCPUTourTwoOpt",
  "MutationCPUPermInversion",
  "MutationCPUPermSwap",
  "SelectionCPUTournament",
]
