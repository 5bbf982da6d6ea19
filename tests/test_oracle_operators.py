"""Pins the GA-operator, RNG and initialisation oracles with hand-worked answers."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import init as oinit
from oracle import operators as ops
from oracle.distances import compute_distance_matrix


def test_tour_cost_known_answers():
  # reference tests/algorithms/strategies/evaluation/test_unit_cost_evaluation.py:21-53
  d = np.array([[0, 2, 9, 10], [2, 0, 6, 4], [9, 6, 0, 3], [10, 4, 3, 0]], dtype=np.int32)
  got = ops.tour_costs(np.array([[0, 1, 2, 3], [0, 2, 1, 3], [1, 3, 2, 0]], dtype=np.int32), d)
  assert got.tolist() == [21.0, 29.0, 18.0]
  assert ops.tour_costs(np.array([[0]], dtype=np.int32), d).tolist() == [0.0]
  assert got.dtype == np.float64


def test_pmx_hand_worked_single_and_chained_mapping():
  p1 = np.array([0, 1, 2, 3, 4, 5, 6, 7], dtype=np.int32)
  p2 = np.array([3, 7, 5, 1, 6, 0, 2, 4], dtype=np.int32)
  assert ops.pmx_pair(p1, p2, 3, 5).tolist() == [1, 7, 0, 3, 4, 5, 2, 6]
  q1 = np.array([0, 1, 2, 3, 4], dtype=np.int32)
  q2 = np.array([1, 2, 0, 4, 3], dtype=np.int32)
  assert ops.pmx_pair(q1, q2, 1, 2).tolist() == [0, 1, 2, 4, 3]
  # cuts given in either order
  assert ops.pmx_pair(q1, q2, 2, 1).tolist() == [0, 1, 2, 4, 3]


def test_pmx_batch_is_permutation_and_keeps_segment():
  rng = np.random.default_rng(4)
  n, p = 50, 9
  pop = np.stack([rng.permutation(n) for _ in range(p)]).astype(np.int32)
  lo, hi = ops.pmx_draw(np.random.default_rng(1), p, n)
  kids = ops.pmx_apply(pop, lo, hi)
  for r in range(p):
    assert sorted(kids[r].tolist()) == list(range(n))
  for k in range(p // 2):
    assert np.array_equal(kids[2 * k, lo[k] : hi[k] + 1], pop[2 * k, lo[k] : hi[k] + 1])
    assert np.array_equal(kids[2 * k + 1, lo[k] : hi[k] + 1], pop[2 * k + 1, lo[k] : hi[k] + 1])
  assert np.array_equal(kids[-1], pop[-1])  # odd row copied


def test_distinct_pairs_draw_order_and_range():
  a = ops.sample_distinct_pairs(np.random.default_rng(0), 1000, 0, 7)
  assert (a[:, 0] != a[:, 1]).all() and a.min() >= 0 and a.max() < 7
  rng = np.random.default_rng(0)
  first = rng.integers(0, 7, size=1000, dtype=np.int32)
  sh = rng.integers(0, 6, size=1000, dtype=np.int32)
  assert np.array_equal(a[:, 0], first) and np.array_equal(a[:, 1], sh + (sh >= first))
  with pytest.raises(ValueError):
    ops.sample_distinct_pairs(np.random.default_rng(0), 3, 0, 1)


def test_inversion_reverses_inclusive_segment_where_prob_below_rate():
  pop = np.tile(np.arange(6, dtype=np.int32), (3, 1))
  out = ops.inversion_apply(pop, np.array([0.01, 0.9, 0.0]), np.array([4, 1, 0]), np.array([1, 3, 5]), 0.05)
  assert out.tolist() == [[0, 4, 3, 2, 1, 5], [0, 1, 2, 3, 4, 5], [5, 4, 3, 2, 1, 0]]
  assert pop.tolist()[0] == [0, 1, 2, 3, 4, 5]


def test_tournament_first_minimum_wins():
  fit = np.array([5.0, 1.0, 1.0, 0.5])
  cand = np.array([[0, 1, 2], [2, 1, 0], [3, 3, 1]], dtype=np.int32)
  assert ops.tournament_winners(cand, fit).tolist() == [1, 2, 3]


def test_survivors_stable_ties_prefer_old_population():
  old = np.array([[0, 1, 2], [1, 2, 0]], dtype=np.int32)
  new = np.array([[2, 1, 0], [0, 2, 1]], dtype=np.int32)
  pop, fit, idx = ops.survivors(old, np.array([3.0, 1.0]), new, np.array([1.0, 0.5]), 3)
  assert idx.tolist() == [3, 1, 2] and fit.tolist() == [0.5, 1.0, 1.0]
  assert pop.tolist() == [[0, 2, 1], [1, 2, 0], [2, 1, 0]]


def test_splitmix64_published_vector():
  # SplitMix64 from state 0: 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F
  s = 0
  outs = []
  for _ in range(3):
    s, z = oinit.splitmix64_next_scalar(s)
    outs.append(z)
  assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
  st, z = oinit.splitmix64_step(np.array([0], dtype=np.uint64))
  assert int(z[0]) == 0xE220A8397B1DCDAF


def test_random_init_matches_scalar_fisher_yates():
  seeds = oinit.draw_row_seeds(np.random.default_rng(42), 5)
  pop = oinit.random_init(np.arange(30), 5, seeds)
  for r in range(5):
    row = list(range(30))
    s = int(seeds[r])
    for i in range(29, 0, -1):
      s, j = oinit.splitmix64_bounded_scalar(s, i + 1)
      row[i], row[j] = row[j], row[i]
    assert pop[r].tolist() == row


def test_nn_init_rows_are_permutations_and_follow_neighbours(berlin52_coords):
  from oracle.distances import compute_candidate_list

  d = compute_distance_matrix(berlin52_coords)
  cand = compute_candidate_list(d, 5)
  seeds = oinit.draw_row_seeds(np.random.default_rng(42), 8)
  pop = oinit.nn_init(np.arange(52), 8, cand, seeds)
  for r in range(8):
    assert sorted(pop[r].tolist()) == list(range(52))
  # most steps go to one of the 5 nearest neighbours
  hits = sum(int(pop[r, s + 1] in cand[pop[r, s]]) for r in range(8) for s in range(51))
  assert hits > 0.7 * 8 * 51


# Hand-worked GRASP nearest-neighbour rows (nearest_neighbor_init.py:31-99).  Each SplitMix64 draw
# is listed as u = (z >> 32) / 2^32, so bounded(b) = floor(u * b); the z values come from the
# published algorithm (Steele, Lea & Flood 2014), restated independently below and checked against
# its published first output for state 0.
_M64 = (1 << 64) - 1


def _sm64_draws(seed, count):
  out, s = [], seed
  for _ in range(count):
    s = (s + 0x9E3779B97F4A7C15) & _M64
    z = ((s ^ (s >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    out.append((z ^ (z >> 31)) & _M64)
  return out


def test_nn_row_hand_worked_shuffle_and_fallback():
  """Seed 7, n = 6, k = 2.  Draws (u): 0.3898 0.0168 0.9008 0.5829 0.4524 0.2494 0.4680 0.3281
  0.1343 0.4131 0.1036.
    start   d0 bounded(6) = 2 -> city 2; unvisited [0 1 5 3 4 | 2]
    step 1  row 2 = [1 3]: i=0 d1 bounded(2) = 0 -> order [0 1], city 1 unvisited -> 1
            unvisited [0 4 5 3 | 1 2]
    step 2  row 1 = [0 2]: i=0 d2 bounded(2) = 1 -> order [1 0], city 2 visited;
            i=1 d3 bounded(1) = 0 -> city 0 unvisited -> 0; unvisited [3 4 5 | 0 1 2]
    step 3  row 0 = [1 2]: i=0 d4 bounded(2) = 0 -> city 1 visited; i=1 d5 bounded(1) = 0 ->
            city 2 visited; fallback d6 bounded(3) = 1 -> unvisited[1] = 4; unvisited [3 5 | 4 ...]
    step 4  row 4 = [3 5]: i=0 d7 bounded(2) = 0 -> city 3 unvisited -> 3; unvisited [5 | 3 ...]
    step 5  row 3 = [2 4]: d8, d9 both visited; fallback d10 bounded(1) = 0 -> 5
  A bound of k instead of k - i at step 2, i = 1 would give j = 2 (outside the row)."""
  draws = _sm64_draws(7, 11)
  assert [round((z >> 32) / 2**32, 4) for z in draws] == [0.3898, 0.0168, 0.9008, 0.5829, 0.4524, 0.2494, 0.468,
                                                          0.3281, 0.1343, 0.4131, 0.1036]
  assert _sm64_draws(0, 1) == [0xE220A8397B1DCDAF]
  cand = np.array([[1, 2], [0, 2], [1, 3], [2, 4], [3, 5], [4, 0]], dtype=np.int32)
  assert oinit.nn_row(np.arange(6, dtype=np.int32), cand, 7).tolist() == [2, 1, 0, 4, 3, 5]


def test_nn_row_hand_worked_swap_and_pop_fallback():
  """Seed 2024, n = 6, k = 1 (every candidate draw is bounded(1) = 0 but still consumed).  Draws (u):
  0.6228 0.0972 0.2986 0.1162 0.8306 0.5524 0.1368 0.5963 0.1689.
    start   d0 bounded(6) = 3 -> city 3; unvisited [0 1 2 5 4 | 3]
    step 1  row 3 = [0]: d1 -> city 0 -> 0; unvisited [4 1 2 5 | 0 3]
    step 2  row 0 = [1]: d2 -> city 1 -> 1; unvisited [4 5 2 | 1 0 3]
    step 3  row 1 = [0]: d3 -> city 0 visited; fallback d4 bounded(3) = 2 -> unvisited[2] = 2
            (in sorted order the pick would be 5: the swap-and-pop order is what is pinned)
    step 4  row 2 = [0]: d5 visited; fallback d6 bounded(2) = 0 -> unvisited[0] = 4
    step 5  row 4 = [0]: d7 visited; fallback d8 bounded(1) = 0 -> 5"""
  draws = _sm64_draws(2024, 9)
  assert [round((z >> 32) / 2**32, 4) for z in draws] == [0.6228, 0.0972, 0.2986, 0.1162, 0.8306, 0.5524, 0.1368,
                                                          0.5963, 0.1689]
  cand = np.array([[1], [0], [0], [0], [0], [0]], dtype=np.int32)
  assert oinit.nn_row(np.arange(6, dtype=np.int32), cand, 2024).tolist() == [3, 0, 1, 2, 4, 5]


def test_nn_row_plausible_mistakes_fail_the_hand_worked_rows(monkeypatch):
  """The two hand-worked rows reject a start draw taken after the first step's shuffle draw and a
  fallback that picks from the sorted unvisited set."""
  cand1 = np.array([[1, 2], [0, 2], [1, 3], [2, 4], [3, 5], [4, 0]], dtype=np.int32)
  cand2 = np.array([[1], [0], [0], [0], [0], [0]], dtype=np.int32)
  good = (oinit.nn_row(np.arange(6, dtype=np.int32), cand1, 7).tolist(),
          oinit.nn_row(np.arange(6, dtype=np.int32), cand2, 2024).tolist())
  assert good == ([2, 1, 0, 4, 3, 5], [3, 0, 1, 2, 4, 5])

  def sorted_fallback_row(customers, candidates, seed):
    # same rules, but the fallback draws an index into the sorted unvisited cities
    n = customers.size
    state, start = oinit.splitmix64_bounded_scalar(seed, n)
    seen = {int(customers[start])}
    out = [int(customers[start])]
    for _ in range(1, n):
      order = list(range(candidates.shape[1]))
      chosen = -1
      for i in range(len(order)):
        state, off = oinit.splitmix64_bounded_scalar(state, len(order) - i)
        order[i], order[i + off] = order[i + off], order[i]
        nb = int(candidates[out[-1], order[i]])
        if nb not in seen:
          chosen = nb
          break
      if chosen < 0:
        left = sorted(set(range(n)) - seen)
        state, u = oinit.splitmix64_bounded_scalar(state, len(left))
        chosen = left[u]
      out.append(chosen)
      seen.add(chosen)
    return out

  assert sorted_fallback_row(np.arange(6), cand2, 2024) != good[1]


def test_greedy_nn_hand_worked():
  from oracle.init import greedy_nn

  x = np.array([0, 1, 3, 6])
  D = np.abs(x[:, None] - x[None, :]).astype(np.int32)
  got = greedy_nn(D, [0, 2, 3])
  assert got.tolist() == [[0, 1, 2, 3], [2, 1, 0, 3], [3, 2, 1, 0]]
  # equal distances: the lower index wins
  x = np.array([-1, 0, 1])
  D = np.abs(x[:, None] - x[None, :]).astype(np.int32)
  assert greedy_nn(D, [1]).tolist() == [[1, 0, 2]]


def test_ox_hand_worked_and_invariants():
  p1 = np.arange(9, dtype=np.int32)
  p2 = np.array([8, 2, 6, 7, 1, 5, 4, 0, 3], dtype=np.int32)
  # Davis' order crossover: segment [3, 5] from p1, the rest in p2's order from position 6 on
  assert ops.ox_pair(p1, p2, 3, 5).tolist() == [6, 7, 1, 3, 4, 5, 0, 8, 2]
  assert ops.ox_pair(p1, p2, 0, 8).tolist() == p1.tolist()  # whole segment: copy of p1
  assert ops.ox_pair(p1, p2, 4, 4).tolist() == [2, 6, 7, 1, 4, 5, 0, 3, 8]
  rng = np.random.default_rng(3)
  pop = np.stack([rng.permutation(50) for _ in range(7)]).astype(np.int32)
  out = ops.ox_batch(np.random.default_rng(1), pop)
  assert (np.sort(out, axis=1) == np.arange(50)).all()
  np.testing.assert_array_equal(out[6], pop[6])  # odd last row copied
  lo, hi = ops.pmx_draw(np.random.default_rng(1), 7, 50)
  for q in range(3):
    c1, c2 = lo[q], hi[q]
    np.testing.assert_array_equal(out[2 * q, c1 : c2 + 1], pop[2 * q, c1 : c2 + 1])
    np.testing.assert_array_equal(out[2 * q + 1, c1 : c2 + 1], pop[2 * q + 1, c1 : c2 + 1])
    # outside the segment the child keeps the other parent's relative order
    seg = set(pop[2 * q, c1 : c2 + 1].tolist())
    rest = [g for g in np.roll(pop[2 * q + 1], -(c2 + 1)) if g not in seg]
    got = [out[2 * q, (c2 + 1 + m) % 50] for m in range(len(rest))]
    assert got == rest


def test_swap_hand_worked():
  pop = np.tile(np.arange(6, dtype=np.int32), (3, 1))
  out = ops.swap_apply(pop, np.array([0.01, 0.9, 0.02]), np.array([0, 1, 5]), np.array([4, 2, 2]), 0.05)
  assert out.tolist() == [[4, 1, 2, 3, 0, 5], [0, 1, 2, 3, 4, 5], [0, 1, 5, 3, 4, 2]]
