"""``EvaluationGPUCost``: HIP tour-cost evaluation (reference src/algorithms/strategies/evaluation/
cost_evaluation.py:165-427, tour_cost.py:32-82) for Hamiltonian TSP tours."""

from __future__ import annotations

from typing import Any

import numpy as np

from .. import _native
from ..protocols import RouteAssessment, tour_assessment
from ..route_helpers import ensure_routes_batch, is_gpu_array, to_device_int32
from ..timing import time_strategy
from .base import StrategyBase


def _prev_power_of_2(v: int, lo: int = 32, hi: int = 256) -> int:
  """Largest power of two <= v, clamped to [lo, hi] (the reference's prev_power_of_2)."""
  v = max(lo, min(hi, int(v)))
  return 1 << (v.bit_length() - 1)


class EvaluationGPUCost(StrategyBase):
  """Constructor of cost_evaluation.py:168-188: `decoder`, `split_mode` and `penalty_policy` are
  the VRP components; for TSP tours they stay at their defaults (no decoder, no penalty), and a
  VRP decoder or penalty policy raises NotImplementedError, as EvaluationCPUCost does."""

  def __init__(self, return_gpu: bool = False, threads_per_block: int = 256, max_matrix_dimension: int | None = None,
               decoder: Any = None, split_mode: str = "prins", penalty_policy: Any = None, *,
               seed: Any = None) -> None:
    if decoder is not None or penalty_policy is not None:
      raise NotImplementedError("EvaluationGPUCost: VRP decoders and penalty policies are not implemented")
    if split_mode not in ("prins", "greedy"):
      raise ValueError(f"split_mode must be 'prins' or 'greedy', got {split_mode!r}")
    super().__init__(seed)
    self.return_gpu = return_gpu
    self.threads_per_block = _prev_power_of_2(threads_per_block)
    self.max_matrix_dimension = max_matrix_dimension
    self.decoder = None
    self.split_mode = split_mode
    self.penalty_policy = None

  def _warmup_impl(self, context) -> None:
    self.evaluate_batch(self._warmup_batch(context, rows=1), context, return_gpu=True)

  def _costs(self, routes, context):
    torch = _native.require_device()
    batch, _ = ensure_routes_batch(routes)
    tours, on_gpu, up = to_device_int32(batch, copy=False)
    self.h2d_bytes += up
    out = torch.empty(tours.shape[0], dtype=torch.float64, device="cuda")
    if tours.shape[0] > 0:
      _native.tour_cost(tours, context.get_gpu_distances(), out, closed=bool(context.return_to_depot))
      self.kernel_launches += 1
    return out, on_gpu

  def _finalize(self, d_costs, on_gpu: bool, return_gpu: bool | None):
    keep = (on_gpu and self.return_gpu) if return_gpu is None else return_gpu
    if keep:
      return d_costs
    host = d_costs.cpu().numpy().astype(np.float64, copy=False)
    self.d2h_bytes += int(host.nbytes)
    return host

  @time_strategy
  def evaluate_batch(self, routes, context, return_gpu: bool | None = None):
    d_costs, on_gpu = self._costs(routes, context)
    return self._finalize(d_costs, on_gpu, return_gpu)

  def evaluate(self, route, context) -> float:
    costs = self.evaluate_batch(route, context, return_gpu=False)
    return float(costs[0])

  def assess_batch(self, routes, context, return_gpu: bool = False) -> RouteAssessment:
    d_costs, on_gpu = self._costs(routes, context)
    return tour_assessment(self._finalize(d_costs, on_gpu, return_gpu))
