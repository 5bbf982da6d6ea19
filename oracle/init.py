"""Oracle: SplitMix64 and the two TSP initialisers (test infrastructure only).

* ``draw_row_seeds`` / ``splitmix64_*`` — src/utils/splitmix64.py:21-56 (SplitMix64 of
  Steele, Lea & Flood, 2014: gamma 0x9E3779B97F4A7C15, mixers 0xBF58476D1CE4E5B9 and
  0x94D049BB133111EB, shifts 30/27/31; bounded draw = ((z >> 32) * bound) >> 32).
* ``random_init``     — InitializationCPUPermRandom.initialize, random_init.py:30-55.
* ``nn_row`` / ``nn_init`` — _nearest_neighbor_row_cpu and InitializationCPUTourNearestNeighbor,
  nearest_neighbor_init.py:31-132.
"""

from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
MUL1 = np.uint64(0xBF58476D1CE4E5B9)
MUL2 = np.uint64(0x94D049BB133111EB)
MASK = 0xFFFFFFFFFFFFFFFF


def draw_row_seeds(rng: np.random.Generator, population_size: int) -> np.ndarray:
  """splitmix64.py:21-25."""
  if population_size <= 0:
    return np.empty(0, dtype=np.uint64)
  return rng.integers(0, np.iinfo(np.int64).max, size=population_size, dtype=np.int64).view(np.uint64)


def splitmix64_step(states: np.ndarray):
  """splitmix64.py:28-35 (vectorised)."""
  with np.errstate(over="ignore"):
    s = states + GAMMA
    z = (s ^ (s >> np.uint64(30))) * MUL1
    z = (z ^ (z >> np.uint64(27))) * MUL2
    z = z ^ (z >> np.uint64(31))
  return s, z


def splitmix64_next_scalar(state: int):
  """splitmix64.py:38-45."""
  s = (state + 0x9E3779B97F4A7C15) & MASK
  z = ((s ^ (s >> 30)) * 0xBF58476D1CE4E5B9) & MASK
  z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
  return s, (z ^ (z >> 31)) & MASK


def splitmix64_bounded_scalar(state: int, bound: int):
  """splitmix64.py:48-51."""
  s, z = splitmix64_next_scalar(state)
  return s, ((z >> 32) * bound) >> 32


def random_init(customers, population_size: int, row_seeds: np.ndarray) -> np.ndarray:
  """Fisher–Yates with one SplitMix64 stream per row, random_init.py:40-55."""
  cust = np.asarray(customers, dtype=np.int32)
  n = int(cust.size)
  if population_size == 0 or n == 0:
    return np.empty((population_size, n), dtype=np.int32)
  pop = np.tile(cust, (population_size, 1)).astype(np.int32, copy=False)
  states = np.asarray(row_seeds, dtype=np.uint64).copy()
  rows = np.arange(population_size, dtype=np.intp)
  for i in range(n - 1, 0, -1):
    states, z = splitmix64_step(states)
    j = (((z >> np.uint64(32)) * np.uint64(i + 1)) >> np.uint64(32)).astype(np.intp)
    tmp = pop[:, i].copy()
    pop[:, i] = pop[rows, j]
    pop[rows, j] = tmp
  return pop


def nn_row(customers: np.ndarray, candidates: np.ndarray, seed_u64: int) -> np.ndarray:
  """GRASP nearest-neighbour tour, nearest_neighbor_init.py:31-99."""
  n = int(customers.size)
  if n == 0:
    return np.empty(0, dtype=np.int32)
  max_node = int(customers.max())
  cand_rows = int(candidates.shape[0])
  k_eff = int(candidates.shape[1]) if candidates.ndim == 2 else 0
  unvisited = customers.astype(np.int32).copy()
  pos_map = np.full(max_node + 1, -1, dtype=np.int32)
  pos_map[unvisited] = np.arange(n, dtype=np.int32)
  state = int(seed_u64)
  rem = n
  out = np.empty(n, dtype=np.int32)
  state, start = splitmix64_bounded_scalar(state, rem)
  current = int(unvisited[start])
  out[0] = current
  last = rem - 1
  last_node = int(unvisited[last])
  unvisited[start] = last_node
  unvisited[last] = current
  pos_map[last_node] = start
  pos_map[current] = last
  rem = last
  order = np.arange(k_eff, dtype=np.int32)
  for step in range(1, n):
    chosen = -1
    if 0 <= current < cand_rows and k_eff > 0:
      order[:] = np.arange(k_eff, dtype=np.int32)
      row = candidates[current]
      for i in range(k_eff):
        state, off = splitmix64_bounded_scalar(state, k_eff - i)
        j = i + off
        order[i], order[j] = order[j], order[i]
        nb = int(row[int(order[i])])
        if 0 <= nb <= max_node:
          p = int(pos_map[nb])
          if 0 <= p < rem:
            chosen = nb
            break
    if chosen < 0:
      state, u = splitmix64_bounded_scalar(state, rem)
      chosen = int(unvisited[u])
    out[step] = chosen
    p = int(pos_map[chosen])
    last = rem - 1
    last_node = int(unvisited[last])
    unvisited[p] = last_node
    unvisited[last] = chosen
    pos_map[last_node] = p
    pos_map[chosen] = last
    rem = last
    current = chosen
  return out


def nn_init(customers, population_size: int, candidates: np.ndarray, row_seeds: np.ndarray) -> np.ndarray:
  """InitializationCPUTourNearestNeighbor.initialize, nearest_neighbor_init.py:109-132."""
  cust = np.asarray(customers, dtype=np.int32)
  pop = np.empty((population_size, cust.size), dtype=np.int32)
  for r in range(population_size):
    pop[r] = nn_row(cust, candidates, int(row_seeds[r]))
  return pop


def greedy_nn(D: np.ndarray, starts) -> np.ndarray:
  """Greedy nearest-neighbour tours: from each start, always the nearest unvisited city, lowest
  index on equal distances.  No reference counterpart (textbook nearest-neighbour heuristic);
  pinned by hand-worked line instances in tests/test_oracle_operators.py."""
  D = np.asarray(D, dtype=np.int64)
  n = D.shape[0]
  out = np.empty((len(starts), n), dtype=np.int32)
  big = np.iinfo(np.int64).max
  for r, s in enumerate(starts):
    seen = np.zeros(n, dtype=bool)
    cur = int(s)
    out[r, 0] = cur
    seen[cur] = True
    for step in range(1, n):
      cur = int(np.argmin(np.where(seen, big, D[cur])))
      out[r, step] = cur
      seen[cur] = True
  return out
