"""StrategyBase: telemetry counters and RNG plumbing shared by every slot
(reference src/algorithms/strategies/base.py:37-155)."""

from __future__ import annotations

import copy
from typing import Any

from ..protocols import RouteAssessment, assess_from_evaluate_batch, canonicalize_default_route
from ..rng import make_rng


class StrategyBase:
  last_elapsed_ms: float
  total_elapsed_ms: float
  call_count: int

  def __init__(self, seed: Any = None) -> None:
    self.seed = seed
    self.rng = make_rng(seed)
    self.h2d_bytes = 0
    self.d2h_bytes = 0
    self.kernel_launches = 0
    self.last_elapsed_ms = 0.0
    self.total_elapsed_ms = 0.0
    self.call_count = 0

  def reset_metrics(self) -> None:
    self.h2d_bytes = 0
    self.d2h_bytes = 0
    self.kernel_launches = 0

  def warmup(self, context) -> None:
    """Launch once without changing counters or either RNG state (base.py:83-118)."""
    saved = (self.h2d_bytes, self.d2h_bytes, self.kernel_launches, self.last_elapsed_ms, self.total_elapsed_ms,
             self.call_count)
    self_rng = copy.deepcopy(self.rng.bit_generator.state) if hasattr(self.rng, "bit_generator") else None
    ctx_h2d = getattr(context, "h2d_bytes", None)
    ctx_rng = getattr(context, "rng", None)
    ctx_state = copy.deepcopy(ctx_rng.bit_generator.state) if hasattr(ctx_rng, "bit_generator") else None
    try:
      self._warmup_impl(context)
    finally:
      (self.h2d_bytes, self.d2h_bytes, self.kernel_launches, self.last_elapsed_ms, self.total_elapsed_ms,
       self.call_count) = saved
      if self_rng is not None:
        self.rng.bit_generator.state = self_rng
      if ctx_h2d is not None:
        context.h2d_bytes = ctx_h2d
      if ctx_state is not None:
        ctx_rng.bit_generator.state = ctx_state

  def _warmup_impl(self, context) -> None:
    del context

  def canonicalize_route(self, route, context):
    del context
    return canonicalize_default_route(route)

  def assess_batch(self, routes, context, return_gpu: bool = False) -> RouteAssessment:
    return assess_from_evaluate_batch(self, routes, context, return_gpu=return_gpu)

  @staticmethod
  def _warmup_batch(context, rows: int = 2):
    """(rows, n) of arange permutations on the device, never zeros (base.py:141-155)."""
    import torch

    return torch.arange(context.dimension, dtype=torch.int32, device="cuda").repeat(rows, 1)
