"""Host random-number plumbing (reference src/utils/rng.py:23-64, splitmix64.py:21-25,
route_helpers.py:74-108).

All stochastic decisions of the GA are drawn here, on the host, from one NumPy ``Generator``
(PCG64) in the reference's order; kernels only consume the drawn buffers.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np


def make_rng(seed=None) -> np.random.Generator:
  """rng.py:23-45."""
  if isinstance(seed, np.random.Generator):
    return seed
  if callable(seed):
    rng = seed()
    if not isinstance(rng, np.random.Generator):
      raise TypeError(f"Callable seed factory must return np.random.Generator, got {type(rng)!r}")
    return rng
  if isinstance(seed, Sequence) and not isinstance(seed, (str, bytes)):
    return np.random.default_rng(np.random.SeedSequence(list(seed)))
  if seed is None or isinstance(seed, (int, np.integer)):
    return np.random.default_rng(seed)
  raise TypeError(f"Unsupported seed type: {type(seed)!r}")


def resolve_rng(context=None, fallback=None) -> np.random.Generator:
  """rng.py:48-64: the context generator wins over the strategy's own."""
  if context is not None and isinstance(getattr(context, "rng", None), np.random.Generator):
    return context.rng
  return make_rng(fallback)


def draw_row_seeds(rng: np.random.Generator, population_size: int) -> np.ndarray:
  """splitmix64.py:21-25: int64 draws viewed as uint64 row seeds."""
  if population_size <= 0:
    return np.empty(0, dtype=np.uint64)
  return rng.integers(0, np.iinfo(np.int64).max, size=population_size, dtype=np.int64).view(np.uint64)


def sample_distinct_pairs(rng: np.random.Generator, num_pairs: int, low: int, high: int) -> np.ndarray:
  """route_helpers.py:74-108: first uniform on [low, high), second = v' + (v' >= first)."""
  if high - low < 2:
    raise ValueError(f"Need at least two distinct integers in [{low}, {high})")
  if num_pairs <= 0:
    return np.empty((0, 2), dtype=np.int32)
  first = rng.integers(low, high, size=num_pairs, dtype=np.int32)
  shifted = rng.integers(low, high - 1, size=num_pairs, dtype=np.int32)
  second = shifted + (shifted >= first).astype(np.int32)
  return np.stack((first, second), axis=1).astype(np.int32, copy=False)
