"""Oracle: the reference CPU 2-opt scan (test infrastructure only).

Restates ``src/algorithms/strategies/local_search/two_opt.py`` of the reference:

* ``exhaustive_pairs``   — ``_exhaustive_pairs``   (two_opt.py:158-201)
* ``pair_blocks``        — ``_pair_blocks``        (two_opt.py:303-367), blocks of 2**18 (:155)
* ``scan_context``       — ``_scan_context``       (two_opt.py:369-384)
* ``pair_deltas``        — ``_pair_deltas``        (two_opt.py:204-261)
* ``step_best``          — ``_step_best``          (two_opt.py:386-440)
* ``step_first``         — ``_step_first``         (two_opt.py:442-489)
* ``step_multi_move``    — ``_step_multi_move``    (two_opt.py:491-563)
* ``refine``             — ``_refine``             (two_opt.py:565-595)
* ``improve_batch``      — ``improve_batch``       (two_opt.py:611-628)

The functions also count the deltas they evaluate (one per admissible pair per pass)
and the moves they apply, which is what the CPU baseline timing reports.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np

PAIR_BLOCK = 1 << 18  # two_opt.py:155


@lru_cache(maxsize=8)
def exhaustive_pairs(n: int, is_closed: bool = False, is_cyclic: bool = True, fixed_start: bool = True):
  """Pair set in scan order, two_opt.py:158-201."""
  i_raw, j_raw = np.triu_indices(n, k=2)
  i_idx = i_raw.astype(np.int32, copy=False)
  j_idx = j_raw.astype(np.int32, copy=False)
  if is_closed:
    keep = j_idx < n - 1
    return i_idx[keep], j_idx[keep]
  if is_cyclic:
    keep = ~((i_idx == 0) & (j_idx == n - 1))
    return i_idx[keep], j_idx[keep]
  if fixed_start:
    return i_idx, j_idx
  prefix_j = np.arange(1, max(1, n - 1), dtype=np.int32)
  prefix_i = np.full(prefix_j.shape, -1, dtype=np.int32)
  return np.concatenate((prefix_i, i_idx)), np.concatenate((prefix_j, j_idx))


def inverse_positions(route: np.ndarray) -> np.ndarray:
  """``route_helpers.py:111-131``: pos[route[idx]] = idx, -1 for absent ids."""
  if route.size == 0:
    return np.full(0, -1, dtype=np.int32)
  size = max(int(route.size), int(route.max()) + 1)
  pos = np.full(size, -1, dtype=np.int32)
  pos[route] = np.arange(route.size, dtype=np.int32)
  return pos


def pair_blocks(tour, positions, candidates, *, is_closed=False, is_cyclic=True, fixed_start=True):
  """Yield (i, j) blocks in scan order, two_opt.py:303-367 (duplicates kept)."""
  n = len(tour)
  if candidates is None or positions is None:
    i_all, j_all = exhaustive_pairs(n, is_closed, is_cyclic, fixed_start)
    for start in range(0, i_all.size, PAIR_BLOCK):
      yield i_all[start : start + PAIR_BLOCK], j_all[start : start + PAIR_BLOCK]
    return
  rows = n - 2
  neighbours = candidates[tour[:rows]]
  in_range = (neighbours >= 0) & (neighbours < positions.size)
  if bool(in_range.all()):
    j_mat = positions[neighbours]
  else:
    j_mat = np.where(in_range, positions[np.clip(neighbours, 0, max(positions.size - 1, 0))], -1)
  i_mat = np.arange(rows, dtype=np.int32)[:, np.newaxis]
  keep = (j_mat > i_mat + 1) & (j_mat < n)
  if rows:
    keep[0, :] &= j_mat[0, :] < n - 1
  i_keep, _ = np.nonzero(keep)
  yield i_keep.astype(np.int32, copy=False), j_mat[keep]


def scan_context(tour, dist, candidates):
  """two_opt.py:369-384."""
  next_node = np.roll(tour, -1)
  edge_costs = dist[tour, next_node].astype(np.int64)
  positions = inverse_positions(tour) if candidates is not None else None
  return next_node, edge_costs, positions


def pair_deltas(tour, next_node, edge_costs, dist, i_idx, j_idx, *, is_closed=False, is_cyclic=True):
  """int64 gain of reversing tour[i+1 : j+1], two_opt.py:204-261."""
  if is_closed or is_cyclic:
    a = tour[i_idx]
    b = next_node[i_idx]
    c = tour[j_idx]
    d = next_node[j_idx]
    return (
      np.asarray(dist[a, c], dtype=np.int64)
      + np.asarray(dist[b, d], dtype=np.int64)
      - edge_costs[i_idx]
      - edge_costs[j_idx]
    )
  n = tour.size
  is_prefix = i_idx < 0
  is_suffix = j_idx == n - 1
  safe_i = np.maximum(i_idx, 0)
  safe_j = np.minimum(j_idx, max(0, n - 2))
  a = tour[safe_i]
  b = np.where(is_prefix, tour[0], next_node[safe_i])
  c = tour[j_idx]
  d = next_node[safe_j]
  d_ac = np.where(is_prefix, np.int64(0), np.asarray(dist[a, c], dtype=np.int64))
  d_bd = np.where(is_suffix, np.int64(0), np.asarray(dist[b, d], dtype=np.int64))
  old_ab = np.where(is_prefix, np.int64(0), edge_costs[safe_i])
  old_cd = np.where(is_suffix, np.int64(0), edge_costs[safe_j])
  return d_ac + d_bd - old_ab - old_cd


class _Counter:
  __slots__ = ("deltas", "moves", "passes")

  def __init__(self) -> None:
    self.deltas = 0
    self.moves = 0
    self.passes = 0


def step_best(tour, dist, candidates, counter=None, **topo) -> bool:
  """Steepest move, first pair in scan order wins a tie, two_opt.py:386-440."""
  n = len(tour)
  if n < 4:
    return False
  next_node, edge_costs, positions = scan_context(tour, dist, candidates)
  best_delta, best_i, best_j = 0, -1, -1
  for i_idx, j_idx in pair_blocks(tour, positions, candidates, **topo):
    if i_idx.size == 0:
      continue
    deltas = pair_deltas(tour, next_node, edge_costs, dist, i_idx, j_idx,
                         is_closed=topo.get("is_closed", False), is_cyclic=topo.get("is_cyclic", True))
    if counter is not None:
      counter.deltas += int(i_idx.size)
    pos = int(np.argmin(deltas))
    gain = int(deltas[pos])
    if gain < best_delta:
      best_delta, best_i, best_j = gain, int(i_idx[pos]), int(j_idx[pos])
  if best_delta < 0:
    tour[best_i + 1 : best_j + 1] = tour[best_i + 1 : best_j + 1][::-1]
    if counter is not None:
      counter.moves += 1
    return True
  return False


def step_first(tour, dist, candidates, counter=None, **topo) -> bool:
  """First improving pair in scan order, two_opt.py:442-489.

  The delta count covers every pair of the blocks evaluated, as the reference
  evaluates whole blocks before testing the mask.
  """
  n = len(tour)
  if n < 4:
    return False
  next_node, edge_costs, positions = scan_context(tour, dist, candidates)
  for i_idx, j_idx in pair_blocks(tour, positions, candidates, **topo):
    if i_idx.size == 0:
      continue
    improving = pair_deltas(tour, next_node, edge_costs, dist, i_idx, j_idx,
                            is_closed=topo.get("is_closed", False), is_cyclic=topo.get("is_cyclic", True)) < 0
    if counter is not None:
      counter.deltas += int(i_idx.size)
    if not bool(improving.any()):
      continue
    pos = int(np.argmax(improving))
    i, j = int(i_idx[pos]), int(j_idx[pos])
    tour[i + 1 : j + 1] = tour[i + 1 : j + 1][::-1]
    if counter is not None:
      counter.moves += 1
    return True
  return False


def step_multi_move(tour, dist, candidates, counter=None, **topo) -> bool:
  """All disjoint improving moves, steepest first (stable), two_opt.py:491-563."""
  n = len(tour)
  if n < 4:
    return False
  next_node, edge_costs, positions = scan_context(tour, dist, candidates)
  gains, starts, ends = [], [], []
  for i_idx, j_idx in pair_blocks(tour, positions, candidates, **topo):
    if i_idx.size == 0:
      continue
    deltas = pair_deltas(tour, next_node, edge_costs, dist, i_idx, j_idx,
                         is_closed=topo.get("is_closed", False), is_cyclic=topo.get("is_cyclic", True))
    if counter is not None:
      counter.deltas += int(i_idx.size)
    improving = deltas < 0
    if not bool(improving.any()):
      continue
    gains.append(deltas[improving])
    starts.append(i_idx[improving])
    ends.append(j_idx[improving])
  if not gains:
    return False
  order = np.argsort(np.concatenate(gains), kind="stable")
  from_idx = np.concatenate(starts)
  to_idx = np.concatenate(ends)
  occupied = np.zeros(n, dtype=bool)
  applied = False
  for slot in order:
    i, j = int(from_idx[slot]), int(to_idx[slot])
    occ_start = max(0, i)
    if occupied[occ_start : j + 1].any():
      continue
    tour[i + 1 : j + 1] = tour[i + 1 : j + 1][::-1]
    occupied[occ_start : j + 1] = True
    applied = True
    if counter is not None:
      counter.moves += 1
  return applied


_STEPS = {"best": step_best, "first": step_first, "multi_move": step_multi_move}


def refine(tour, dist, candidates, mode="best", max_iterations=10, counter=None, **topo) -> np.ndarray:
  """Up to ``max_iterations`` acceptance steps on a copy, two_opt.py:565-595."""
  improved = np.array(tour, dtype=np.int32, copy=True)
  if len(improved) < 4 or max_iterations <= 0:
    return improved
  step = _STEPS[mode]
  for _ in range(max_iterations):
    if counter is not None:
      counter.passes += 1
    if not step(improved, dist, candidates, counter, **topo):
      break
  return improved


def improve_batch(routes, dist, candidates=None, mode="best", max_iterations=10, counter=None, **topo):
  """Serial loop over rows, two_opt.py:611-628 (candidates already resolved)."""
  batch = np.asarray(routes, dtype=np.int32)
  was_1d = batch.ndim == 1
  batch = batch.reshape(1, -1) if was_1d else batch
  out = np.empty_like(batch)
  for r in range(batch.shape[0]):
    out[r] = refine(batch[r], dist, candidates, mode, max_iterations, counter, **topo)
  return out[0] if was_1d else out


def new_counter() -> _Counter:
  return _Counter()
