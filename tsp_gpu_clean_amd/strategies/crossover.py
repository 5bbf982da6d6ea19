"""Pair crossovers on the device: ``CrossoverGPUPermPartiallyMapped`` (reference
src/algorithms/strategies/crossover/pmx_crossover.py:174-317) and ``CrossoverGPUPermOrder``
(order_crossover.py:179-285).  Both draw their cut pairs on the host exactly as the reference
strategies do (one ``sample_distinct_pairs`` call per batch, pairs (2q, 2q+1) sharing a cut)."""

from __future__ import annotations

from typing import Any

import numpy as np

from .. import _native
from ..route_helpers import ensure_routes_batch, is_gpu_array, resolve_route_bounds, restore_routes_shape, to_device_int32
from ..rng import resolve_rng, sample_distinct_pairs
from ..timing import time_strategy
from .base import StrategyBase


class _PairCutCrossover(StrategyBase):
  """Host-drawn cuts, one kernel launch per batch; subclasses name the kernel."""

  _name = "crossover"

  def __init__(self, seed: Any = None) -> None:
    super().__init__(seed)

  def _kernel(self, parents, d_lo, d_hi, out) -> None:
    raise NotImplementedError

  def _warmup_impl(self, context) -> None:
    self.crossover_batch(self._warmup_batch(context, rows=2), context=context)

  def _run(self, parents, lo: np.ndarray, hi: np.ndarray):
    torch = _native.require_device()
    d_lo = torch.from_numpy(np.ascontiguousarray(lo, dtype=np.int32)).to("cuda")
    d_hi = torch.from_numpy(np.ascontiguousarray(hi, dtype=np.int32)).to("cuda")
    self.h2d_bytes += int(lo.nbytes + hi.nbytes)
    out = torch.empty_like(parents)
    self._kernel(parents, d_lo, d_hi, out)
    self.kernel_launches += 1
    return out

  def crossover(self, parent1, parent2, context=None):
    """One child of (parent1, parent2) with one host-drawn cut pair (pmx_crossover.py:107-126,
    order_crossover.py:97-114)."""
    torch = _native.require_device()
    on_gpu = is_gpu_array(parent1)
    p1 = parent1 if on_gpu else torch.as_tensor(np.asarray(parent1, dtype=np.int32))
    p2 = parent2 if is_gpu_array(parent2) else torch.as_tensor(np.asarray(parent2, dtype=np.int32))
    n = int(p1.shape[0])
    if n <= 1:
      return p1.clone() if on_gpu else np.asarray(parent1, dtype=np.int32).copy()
    rng = resolve_rng(context, self.rng)
    low, high = resolve_route_bounds(n, context=context, route_sample=parent1)
    if high - low < 2:
      return p1.clone() if on_gpu else np.asarray(parent1, dtype=np.int32).copy()
    if (low, high) != (0, n):
      raise NotImplementedError(f"depot-bounded routes are not supported by the device {self._name}")
    cuts = sample_distinct_pairs(rng, 1, low, high)[0]
    pair = torch.stack([p1.to("cuda", torch.int32), p2.to("cuda", torch.int32)]).contiguous()
    child = self._run(pair, np.array([min(cuts)]), np.array([max(cuts)]))[0]
    return child if on_gpu else child.cpu().numpy()

  @time_strategy
  def crossover_batch(self, routes, context=None):
    """Pairs (2q, 2q+1) share host-drawn cuts; an odd last row is copied (pmx_crossover.py:128-171,
    order_crossover.py:116-176)."""
    _native.require_device()
    batch, was_1d = ensure_routes_batch(routes)
    num, n = int(batch.shape[0]), int(batch.shape[1])
    on_gpu = is_gpu_array(batch)
    if was_1d or num < 2:
      return restore_routes_shape(batch.clone() if on_gpu else batch.copy(), was_1d)
    low, high = resolve_route_bounds(n, context=context, route_sample=batch[0])
    if n <= 1 or high - low < 2:
      return batch.clone() if on_gpu else batch.copy()
    if (low, high) != (0, n):
      raise NotImplementedError(f"depot-bounded routes are not supported by the device {self._name}")
    rng = resolve_rng(context, self.rng)
    pairs = sample_distinct_pairs(rng, num // 2, low, high)
    lo = np.minimum(pairs[:, 0], pairs[:, 1]).astype(np.int32)
    hi = np.maximum(pairs[:, 0], pairs[:, 1]).astype(np.int32)
    parents, _, up = to_device_int32(batch, copy=False)
    self.h2d_bytes += up
    out = self._run(parents, lo, hi)
    if on_gpu:
      return out
    host = out.cpu().numpy()
    self.d2h_bytes += int(host.nbytes)
    return host


class CrossoverGPUPermPartiallyMapped(_PairCutCrossover):
  _name = "PMX"

  def _kernel(self, parents, d_lo, d_hi, out) -> None:
    _native.pmx(parents, d_lo, d_hi, out)


class CrossoverGPUPermOrder(_PairCutCrossover):
  """Order crossover (OX): the child keeps its parent's segment and fills the rest with the other
  parent's genes in that parent's order, starting after the segment (order_crossover.py:49-81)."""

  _name = "OX"

  def _kernel(self, parents, d_lo, d_hi, out) -> None:
    _native.ox(parents, d_lo, d_hi, out)
