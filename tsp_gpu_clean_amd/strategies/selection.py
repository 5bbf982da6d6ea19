"""``SelectionGPUTournament`` (reference src/algorithms/strategies/selection/tournament_selection.py:131-248)."""

from __future__ import annotations

from typing import Any

import numpy as np

from .. import _native
from ..route_helpers import ensure_routes_batch, is_gpu_array, to_device_float64, to_device_int32
from ..rng import resolve_rng
from ..timing import time_strategy
from .base import StrategyBase


class SelectionGPUTournament(StrategyBase):
  def __init__(self, k: int = 5, seed: Any = None) -> None:
    if k < 1:
      raise ValueError(f"Tournament size k must be >= 1, got {k}")
    super().__init__(seed)
    self.k = k

  def _warmup_impl(self, context) -> None:
    import torch

    pop = self._warmup_batch(context, rows=2)
    self.select_batch(pop, torch.zeros(2, dtype=torch.float64, device="cuda"), 2, context)

  def select(self, population, fitness, context=None):
    """Single parent: k distinct competitors, first minimum wins (tournament_selection.py:50-76)."""
    batch, _ = ensure_routes_batch(population)
    fit = fitness.detach().cpu().numpy() if is_gpu_array(fitness) else np.asarray(fitness, dtype=np.float64)
    size = int(batch.shape[0])
    rng = resolve_rng(context, self.rng)
    idx = rng.choice(size, size=min(self.k, size), replace=False)
    w = int(idx[int(np.argmin(fit[idx]))])
    return batch[w].clone() if is_gpu_array(batch) else batch[w].copy()

  @time_strategy
  def select_batch(self, routes, fitness, num_parents: int | None = None, context=None):
    """(num_parents, n) winners; host-drawn competitors (tournament_selection.py:112-123)."""
    torch = _native.require_device()
    batch, _ = ensure_routes_batch(routes)
    size = int(batch.shape[0])
    n_select = size if num_parents is None else int(num_parents)
    rng = resolve_rng(context, self.rng)
    cand_np = rng.integers(0, size, size=(n_select, self.k), dtype=np.int32)
    pop, on_gpu, up = to_device_int32(batch, copy=False)
    fit, upf = to_device_float64(fitness)
    cand = torch.from_numpy(cand_np).to("cuda")
    self.h2d_bytes += up + upf + int(cand_np.nbytes)
    out = torch.empty((n_select, pop.shape[1]), dtype=torch.int32, device="cuda")
    _native.tournament(pop, fit, cand, out)
    self.kernel_launches += 1
    if on_gpu:
      return out
    host = out.cpu().numpy()
    self.d2h_bytes += int(host.nbytes)
    return host
