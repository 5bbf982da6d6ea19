"""Position-pair mutations on the device: ``MutationGPUPermInversion`` (reference
src/algorithms/strategies/mutation/inversion_mutation.py:102-214) and ``MutationGPUPermSwap``
(swap_mutation.py:108-215).  Both draw probs = random(N) first, then N position pairs, from the
same generator and in the same order as the reference strategies."""

from __future__ import annotations

from typing import Any

import numpy as np

from .. import _native
from ..route_helpers import ensure_routes_batch, is_gpu_array, resolve_route_bounds, restore_routes_shape, to_device_int32
from ..rng import resolve_rng, sample_distinct_pairs
from ..timing import time_strategy
from .base import StrategyBase


class _PositionPairMutation(StrategyBase):
  _name = "mutation"

  def __init__(self, seed: Any = None, candidate_k: int | None = None) -> None:
    super().__init__(seed)
    if candidate_k is not None and candidate_k > 0:
      raise NotImplementedError(f"candidate-guided {self._name} positions are not implemented on the device")
    self.candidate_k = candidate_k

  def _kernel(self, d_routes, d_probs, d_p1, d_p2, rate: float) -> None:
    raise NotImplementedError

  def _warmup_impl(self, context) -> None:
    self.mutate_batch(self._warmup_batch(context, rows=1), 1.0, context=context)

  def mutate(self, individual, mutation_rate: float, context=None):
    return self.mutate_batch(individual, mutation_rate, context=context)

  @time_strategy
  def mutate_batch(self, routes, mutation_rate: float, context=None):
    """probs = random(N) first, then N position pairs; the kernel acts on rows with probs < rate."""
    torch = _native.require_device()
    batch, was_1d = ensure_routes_batch(routes)
    num, n = int(batch.shape[0]), int(batch.shape[1])
    on_gpu = is_gpu_array(batch)
    low, high = resolve_route_bounds(n, context=context, route_sample=batch[0])
    if high - low < 2:
      return restore_routes_shape(batch.clone() if on_gpu else batch.copy(), was_1d)
    if (low, high) != (0, n):
      raise NotImplementedError(f"depot-bounded routes are not supported by the device {self._name}")
    rng = resolve_rng(context, self.rng)
    probs = rng.random(num)
    pairs = sample_distinct_pairs(rng, num, low, high)
    d_routes, _, up = to_device_int32(batch, copy=True)
    d_probs = torch.from_numpy(probs).to("cuda")
    d_p1 = torch.from_numpy(np.ascontiguousarray(pairs[:, 0])).to("cuda")
    d_p2 = torch.from_numpy(np.ascontiguousarray(pairs[:, 1])).to("cuda")
    self.h2d_bytes += up + int(probs.nbytes + pairs.nbytes)
    self._kernel(d_routes, d_probs, d_p1, d_p2, float(mutation_rate))
    self.kernel_launches += 1
    if on_gpu:
      return restore_routes_shape(d_routes, was_1d)
    host = d_routes.cpu().numpy()
    self.d2h_bytes += int(host.nbytes)
    return restore_routes_shape(host, was_1d)


class MutationGPUPermInversion(_PositionPairMutation):
  """Reverse positions [min, max] of the drawn pair (inversion_mutation.py:94-99)."""

  _name = "inversion"

  def _kernel(self, d_routes, d_probs, d_p1, d_p2, rate: float) -> None:
    _native.inversion(d_routes, d_probs, d_p1, d_p2, rate)


class MutationGPUPermSwap(_PositionPairMutation):
  """Exchange the two drawn positions (swap_mutation.py:84-105)."""

  _name = "swap"

  def _kernel(self, d_routes, d_probs, d_p1, d_p2, rate: float) -> None:
    _native.swap(d_routes, d_probs, d_p1, d_p2, rate)
