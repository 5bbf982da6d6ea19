"""Device initialisers ``InitializationGPUPermRandom`` / ``InitializationGPUTourNearestNeighbor``
(reference src/algorithms/strategies/initialization/random_init.py:58-118,
nearest_neighbor_init.py:135-238)."""

from __future__ import annotations

from typing import Any

import numpy as np

from .. import _native
from ..rng import draw_row_seeds, resolve_rng
from ..timing import time_strategy
from .base import StrategyBase


def _validate_k(k: int) -> int:
  if isinstance(k, bool) or not isinstance(k, int) or k < 1:
    raise ValueError(f"Candidate-list width k must be >= 1, got {k!r}")
  return k


class InitializationGPUPermRandom(StrategyBase):
  def __init__(self, seed: Any = None, return_gpu: bool = True) -> None:
    super().__init__(seed)
    self.return_gpu = return_gpu

  def _warmup_impl(self, context) -> None:
    self.initialize(list(range(max(1, context.dimension))), 2, context)

  @time_strategy
  def initialize(self, customers, population_size: int, context=None):
    if population_size < 0:
      raise ValueError(f"population_size must be >= 0, got {population_size}")
    torch = _native.require_device()
    cust = np.ascontiguousarray(customers, dtype=np.int32)
    seeds = draw_row_seeds(resolve_rng(context, self.rng), population_size)
    out = torch.empty((population_size, cust.size), dtype=torch.int32, device="cuda")
    if population_size > 0 and cust.size > 0:
      d_seeds = torch.from_numpy(seeds.view(np.int64)).to("cuda")
      d_cust = torch.from_numpy(cust).to("cuda")
      self.h2d_bytes += int(seeds.nbytes + cust.nbytes)
      _native.random_init(d_seeds, d_cust, out)
      self.kernel_launches += 1
    if self.return_gpu:
      return out
    host = out.cpu().numpy()
    self.d2h_bytes += int(host.nbytes)
    return host


class InitializationGPUTourNearestNeighbor(StrategyBase):
  def __init__(self, k: int = 20, seed: Any = None, return_gpu: bool = True) -> None:
    super().__init__(seed)
    self.k = _validate_k(k)
    self.return_gpu = return_gpu

  def _warmup_impl(self, context) -> None:
    if context.dimension > 1:
      self.initialize(list(range(context.dimension)), 2, context)

  @time_strategy
  def initialize(self, customers, population_size: int, context=None):
    if context is None:
      raise ValueError("InitializationGPUTourNearestNeighbor requires a ProblemContext to read the k-NN candidate list")
    if population_size < 0:
      raise ValueError(f"population_size must be >= 0, got {population_size}")
    torch = _native.require_device()
    cust = np.ascontiguousarray(customers, dtype=np.int32)
    n = int(cust.size)
    seeds = draw_row_seeds(resolve_rng(context, self.rng), population_size)
    out = torch.empty((population_size, n), dtype=torch.int32, device="cuda")
    if population_size > 0 and n > 0:
      if n != context.dimension or int(cust.max()) >= n or int(cust.min()) < 0:
        raise NotImplementedError("device NN init needs customers to be a permutation of 0..dimension-1")
      cand = context.get_gpu_candidate_list(self.k)
      d_seeds = torch.from_numpy(seeds.view(np.int64)).to("cuda")
      d_cust = torch.from_numpy(cust).to("cuda")
      work = torch.empty(2 * population_size * n, dtype=torch.int32, device="cuda")
      self.h2d_bytes += int(seeds.nbytes + cust.nbytes)
      _native.nn_init(d_seeds, d_cust, cand if cand.shape[1] > 0 else None, work, out)
      self.kernel_launches += 1
    if self.return_gpu:
      return out
    host = out.cpu().numpy()
    self.d2h_bytes += int(host.nbytes)
    return host
