"""Oracle: GA operators on host NumPy (test infrastructure only).

Restates, for the TSP permutation case (open permutations, ``resolve_route_bounds``
returns ``(0, n)``, route_helpers.py:30-71):

* ``sample_distinct_pairs``   — route_helpers.py:74-108
* ``tour_costs``              — evaluate_tour_cpu, tour_cost.py:17-29
* ``tournament_select``       — SelectionCPUTournament.select_batch, tournament_selection.py:78-128
* ``pmx_pair`` / ``pmx_batch`` — CrossoverCPUPermPartiallyMapped._apply_pmx / crossover_batch,
                                 pmx_crossover.py:49-105, :128-171
* ``inversion_batch``         — MutationCPUPermInversion.mutate_batch, inversion_mutation.py:65-99
* ``survivors``               — GeneticAlgorithm._select_survivors_assessed for feasible
                                 TSP assessments, genetic_algorithm.py:618-706 (lexsort :689)
* ``elite_indices``           — EliteOnlyLocalSearchStep.run, pipeline_steps.py:153-163
"""

from __future__ import annotations

import numpy as np


def sample_distinct_pairs(rng: np.random.Generator, num_pairs: int, low: int, high: int) -> np.ndarray:
  """route_helpers.py:74-108: first uniform, second = shifted + (shifted >= first)."""
  if high - low < 2:
    raise ValueError(f"Need at least two distinct integers in [{low}, {high})")
  if num_pairs <= 0:
    return np.empty((0, 2), dtype=np.int32)
  first = rng.integers(low, high, size=num_pairs, dtype=np.int32)
  shifted = rng.integers(low, high - 1, size=num_pairs, dtype=np.int32)
  second = shifted + (shifted >= first).astype(np.int32)
  return np.stack((first, second), axis=1).astype(np.int32, copy=False)


def tour_costs(batch: np.ndarray, dist: np.ndarray, return_to_depot: bool = True) -> np.ndarray:
  """float64 closed-tour costs, tour_cost.py:17-29."""
  batch = np.asarray(batch, dtype=np.int32)
  if batch.ndim == 1:
    batch = batch.reshape(1, -1)
  p, n = batch.shape
  if p == 0 or n <= 1:
    return np.zeros(p, dtype=np.float64)
  flat = dist.ravel()
  stride = dist.shape[1]
  costs = flat[batch[:, :-1].astype(np.int64) * stride + batch[:, 1:]].sum(axis=1, dtype=np.float64)
  if return_to_depot:
    costs += flat[batch[:, -1].astype(np.int64) * stride + batch[:, 0]]
  return costs


def tournament_draw(rng: np.random.Generator, pop_size: int, n_select: int, k: int) -> np.ndarray:
  """Candidate draw of tournament_selection.py:113."""
  return rng.integers(0, pop_size, size=(n_select, k), dtype=np.int32)


def tournament_winners(candidates: np.ndarray, fitness: np.ndarray) -> np.ndarray:
  """Winner index per row, first minimum wins, tournament_selection.py:125-127."""
  fit = np.asarray(fitness, dtype=np.float64)
  cols = np.argmin(fit[candidates], axis=1)
  return candidates[np.arange(candidates.shape[0]), cols]


def tournament_select(rng, population, fitness, k, num_parents=None):
  """SelectionCPUTournament.select_batch, tournament_selection.py:78-128."""
  pop = np.asarray(population, dtype=np.int32)
  n_select = pop.shape[0] if num_parents is None else num_parents
  cand = tournament_draw(rng, pop.shape[0], n_select, k)
  return pop[tournament_winners(cand, fitness)].copy()


def pmx_pair(p1: np.ndarray, p2: np.ndarray, c1: int, c2: int) -> np.ndarray:
  """One PMX child for an open permutation, pmx_crossover.py:49-105 (is_closed False)."""
  n = len(p1)
  if n <= 1:
    return p1.copy()
  if c1 > c2:
    c1, c2 = c2, c1
  child = np.full(n, -1, dtype=np.int32)
  child[c1 : c2 + 1] = p1[c1 : c2 + 1]
  max_node = max(n, int(p1.max()) + 1, int(p2.max()) + 1)
  inv_p2 = np.full(max_node, -1, dtype=np.int32)
  inv_p2[p2] = np.arange(n, dtype=np.int32)
  in_seg = np.zeros(max_node, dtype=bool)
  in_seg[p1[c1 : c2 + 1]] = True
  seg_len = c2 - c1 + 1
  for idx in range(c1, c2 + 1):
    val = int(p2[idx])
    if val >= 0 and not in_seg[val]:
      pos, steps = idx, 0
      while c1 <= pos <= c2 and steps <= seg_len:
        mapped = int(p1[pos])
        if mapped < 0:
          break
        nxt = int(inv_p2[mapped])
        if nxt < 0:
          break
        pos = nxt
        steps += 1
      if 0 <= pos < n and (pos < c1 or pos > c2):
        child[pos] = val
  unmapped = child == -1
  child[unmapped] = p2[unmapped]
  return child


def pmx_draw(rng: np.random.Generator, num_routes: int, n: int) -> tuple[np.ndarray, np.ndarray]:
  """Ordered cut pairs for ``num_routes // 2`` pairs, pmx_crossover.py:153-160."""
  pairs = sample_distinct_pairs(rng, num_routes // 2, 0, n)
  return (np.minimum(pairs[:, 0], pairs[:, 1]).astype(np.int32),
          np.maximum(pairs[:, 0], pairs[:, 1]).astype(np.int32))


def pmx_apply(batch: np.ndarray, cut_lo: np.ndarray, cut_hi: np.ndarray) -> np.ndarray:
  """Children (2k, 2k+1) from parents (2k, 2k+1), odd last row copied, pmx_crossover.py:157-169."""
  batch = np.asarray(batch, dtype=np.int32)
  out = np.empty_like(batch)
  for k in range(batch.shape[0] // 2):
    i = 2 * k
    out[i] = pmx_pair(batch[i], batch[i + 1], int(cut_lo[k]), int(cut_hi[k]))
    out[i + 1] = pmx_pair(batch[i + 1], batch[i], int(cut_lo[k]), int(cut_hi[k]))
  if batch.shape[0] % 2 == 1:
    out[-1] = batch[-1]
  return out


def pmx_batch(rng, batch):
  """CrossoverCPUPermPartiallyMapped.crossover_batch, pmx_crossover.py:128-171."""
  batch = np.asarray(batch, dtype=np.int32)
  if batch.shape[0] < 2 or batch.shape[1] < 2:
    return batch.copy()
  lo, hi = pmx_draw(rng, batch.shape[0], batch.shape[1])
  return pmx_apply(batch, lo, hi)


def inversion_draw(rng: np.random.Generator, num_routes: int, n: int):
  """probs first, then the position pairs, inversion_mutation.py:84-92."""
  probs = rng.random(num_routes)
  pairs = sample_distinct_pairs(rng, num_routes, 0, n)
  return probs, pairs[:, 0].copy(), pairs[:, 1].copy()


def inversion_apply(batch, probs, pos1, pos2, rate):
  """Reverse [min, max] inclusive where probs < rate, inversion_mutation.py:94-99."""
  out = np.array(batch, dtype=np.int32, copy=True)
  for i in np.flatnonzero(probs < rate):
    s, e = int(min(pos1[i], pos2[i])), int(max(pos1[i], pos2[i]))
    out[i, s : e + 1] = out[i, s : e + 1][::-1]
  return out


def inversion_batch(rng, batch, rate):
  """MutationCPUPermInversion.mutate_batch, inversion_mutation.py:65-99."""
  batch = np.asarray(batch, dtype=np.int32)
  if batch.shape[1] < 2:
    return batch.copy()
  probs, p1, p2 = inversion_draw(rng, batch.shape[0], batch.shape[1])
  return inversion_apply(batch, probs, p1, p2, rate)


def ox_pair(p1: np.ndarray, p2: np.ndarray, c1: int, c2: int) -> np.ndarray:
  """Order crossover of one child, open permutation, order_crossover.py:49-81 (is_closed False):
  child[c1..c2] = p1[c1..c2]; the other positions, from c2+1 onward (wrapping), take p2's
  genes outside that segment in p2's order starting from position c2+1 (wrapping)."""
  n = len(p1)
  child = np.full(n, -1, dtype=np.int32)
  child[c1 : c2 + 1] = p1[c1 : c2 + 1]
  in_seg = np.zeros(max(n, int(p1.max()) + 1, int(p2.max()) + 1), dtype=bool)
  in_seg[p1[c1 : c2 + 1]] = True
  shift = (c2 + 1) % n
  p2_rot = np.roll(p2, -shift)
  remaining = p2_rot[~in_seg[p2_rot]]
  rot_pos = (np.arange(n) + shift) % n
  empty = child[rot_pos] == -1
  child[rot_pos[empty]] = remaining
  return child


def ox_apply(batch: np.ndarray, cut_lo: np.ndarray, cut_hi: np.ndarray) -> np.ndarray:
  """Pairs (2k, 2k+1) share cuts; odd last row copied, order_crossover.py:129-176."""
  batch = np.asarray(batch, dtype=np.int32)
  out = batch.copy()
  for q in range(batch.shape[0] // 2):
    i = 2 * q
    out[i] = ox_pair(batch[i], batch[i + 1], int(cut_lo[q]), int(cut_hi[q]))
    out[i + 1] = ox_pair(batch[i + 1], batch[i], int(cut_lo[q]), int(cut_hi[q]))
  return out


def ox_batch(rng, batch):
  """CrossoverCPUPermOrder.crossover_batch, order_crossover.py:129-176 (same draws as PMX)."""
  batch = np.asarray(batch, dtype=np.int32)
  if batch.shape[0] < 2 or batch.shape[1] < 2:
    return batch.copy()
  lo, hi = pmx_draw(rng, batch.shape[0], batch.shape[1])
  return ox_apply(batch, lo, hi)


def swap_apply(batch, probs, pos1, pos2, rate):
  """Swap positions pos1/pos2 of the rows with probs < rate, swap_mutation.py:84-105."""
  out = np.array(batch, dtype=np.int32, copy=True)
  act = np.flatnonzero(probs < rate)
  out[act, pos1[act]], out[act, pos2[act]] = out[act, pos2[act]], out[act, pos1[act]]
  return out


def swap_batch(rng, batch, rate):
  """MutationCPUPermSwap.mutate_batch without candidate lists: probs, then the position pairs."""
  batch = np.asarray(batch, dtype=np.int32)
  if batch.shape[1] < 2:
    return batch.copy()
  probs, p1, p2 = inversion_draw(rng, batch.shape[0], batch.shape[1])
  return swap_apply(batch, probs, p1, p2, rate)


def survivors(old_pop, old_fit, new_pop, new_fit, pop_size):
  """(mu + lambda) survival with all-feasible, zero-violation TSP assessments.

  ``lexsort((obj, norm_viol, ~feas))`` (genetic_algorithm.py:682-689) reduces to a
  stable ascending sort on the concatenated objectives.
  """
  combined_pop = np.vstack([old_pop, new_pop])
  obj = np.concatenate([np.asarray(old_fit, np.float64), np.asarray(new_fit, np.float64)])
  zeros = np.zeros_like(obj)
  idx = np.lexsort(np.vstack([obj, zeros, zeros]))[:pop_size]
  return combined_pop[idx], obj[idx], idx


def elite_indices(fitness: np.ndarray, k: int) -> np.ndarray:
  """``argsort(fitness)[:k]``, pipeline_steps.py:163 (default kind)."""
  return np.argsort(np.asarray(fitness, dtype=np.float64))[:k]
