"""Pins the 2-opt oracle (NumPy and C restatements) against independent checks.

* The nested-loop scalar specification below restates the reference's own test spec
  (tests/algorithms/strategies/local_search/test_unit_two_opt_vectorized_scan.py:38-204
  of the reference): sequential loops, strict ``<`` for ``best``, first improving pair
  for ``first``, sort-then-greedy for ``multi_move``.
* Pair-order and duplicate-candidate known answers (same file :311-350).
* Per-move identity cost(after) = cost(before) + delta and brute-force local optimality.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_two_opt
from oracle import two_opt as o2
from oracle.operators import tour_costs
from tests.conftest import lattice_matrix, nearest, rounded_matrix


def _spec_pairs(n, tour, dist, cand):
  if cand is None:
    for i in range(n - 2):
      limit = n - 1 if i == 0 else n
      for j in range(i + 2, limit):
        yield i, j
    return
  pos = np.full(n, -1)
  pos[tour] = np.arange(n)
  for i in range(n - 2):
    limit = n - 1 if i == 0 else n
    for r in range(cand.shape[1]):
      nb = int(cand[int(tour[i]), r])
      if nb < 0 or nb >= n:
        continue
      j = int(pos[nb])
      if j <= i + 1 or j >= limit:
        continue
      yield i, j


def _spec_delta(tour, dist, i, j):
  n = len(tour)
  a, b, c, d = int(tour[i]), int(tour[i + 1]), int(tour[j]), int(tour[(j + 1) % n])
  return int(dist[a, c]) + int(dist[b, d]) - int(dist[a, b]) - int(dist[c, d])


def _spec_step(tour, dist, cand, mode):
  n = len(tour)
  pairs = list(_spec_pairs(n, tour, dist, cand))
  if mode == "best":
    best, bi, bj = 0, -1, -1
    for i, j in pairs:
      dl = _spec_delta(tour, dist, i, j)
      if dl < best:
        best, bi, bj = dl, i, j
    if best < 0:
      tour[bi + 1 : bj + 1] = tour[bi + 1 : bj + 1][::-1]
      return True
    return False
  if mode == "first":
    for i, j in pairs:
      if _spec_delta(tour, dist, i, j) < 0:
        tour[i + 1 : j + 1] = tour[i + 1 : j + 1][::-1]
        return True
    return False
  moves = [(_spec_delta(tour, dist, i, j), i, j) for i, j in pairs]
  moves = [m for m in moves if m[0] < 0]
  if not moves:
    return False
  moves.sort(key=lambda m: m[0])
  applied = []
  for _, i, j in moves:
    if any(not (j < s or i > e) for s, e in applied):
      continue
    tour[i + 1 : j + 1] = tour[i + 1 : j + 1][::-1]
    applied.append((i, j))
  return True


def _spec_refine(tour, dist, cand, mode, iters):
  t = tour.copy()
  if len(t) < 4 or iters <= 0:
    return t
  for _ in range(iters):
    if not _spec_step(t, dist, cand, mode):
      break
  return t


@pytest.mark.parametrize("mode", ["best", "first", "multi_move"])
@pytest.mark.parametrize("k", [None, 3, 20])
@pytest.mark.parametrize("n", [4, 5, 6, 8, 10, 17, 40])
def test_oracle_matches_scalar_spec(mode, k, n):
  dist = rounded_matrix(n, seed=n)
  cand = None if k is None else nearest(dist, k)
  rng = np.random.default_rng(n * 100 + (k or 0))
  for _ in range(3):
    start = rng.permutation(n).astype(np.int32)
    for iters in (1, 3, 10):
      expected = _spec_refine(start, dist, cand, mode, iters)
      got = o2.refine(start, dist, cand, mode, iters)
      np.testing.assert_array_equal(got, expected)


@pytest.mark.parametrize("mode", ["best", "first", "multi_move"])
@pytest.mark.parametrize("k", [None, 3])
@pytest.mark.parametrize("n", [9, 16, 25])
def test_oracle_tie_breaking_on_lattice(mode, k, n):
  dist = lattice_matrix(n)
  cand = None if k is None else nearest(dist, k)
  for trial in range(4):
    start = np.random.default_rng(n * 31 + trial).permutation(n).astype(np.int32)
    np.testing.assert_array_equal(o2.refine(start, dist, cand, mode, 10), _spec_refine(start, dist, cand, mode, 10))


@pytest.mark.parametrize("n", [4, 5, 6, 7, 17, 40])
def test_exhaustive_pair_order(n):
  i_idx, j_idx = o2.exhaustive_pairs(n)
  expected = [(i, j) for i in range(n - 2) for j in range(i + 2, n - 1 if i == 0 else n)]
  assert list(zip(i_idx.tolist(), j_idx.tolist())) == expected
  assert len(expected) == n * (n - 3) // 2


def test_candidate_duplicates_kept():
  n = 8
  tour = np.arange(n, dtype=np.int32)
  cand = np.tile(np.array([3, 5], dtype=np.int32), (n, 1))
  cand[0] = [3, 3]
  blocks = list(o2.pair_blocks(tour, o2.inverse_positions(tour), cand))
  i_all = np.concatenate([b[0] for b in blocks])
  j_all = np.concatenate([b[1] for b in blocks])
  assert int(np.count_nonzero((i_all == 0) & (j_all == 3))) == 2


@pytest.mark.parametrize("k", [None, 5])
def test_moves_change_cost_by_delta_and_converge_to_local_optimum(k):
  n = 60
  dist = rounded_matrix(n, seed=7)
  cand = None if k is None else nearest(dist, k)
  t = np.random.default_rng(3).permutation(n).astype(np.int32)
  for _ in range(400):
    before = tour_costs(t, dist)[0]
    pairs = list(_spec_pairs(n, t, dist, cand))
    deltas = [_spec_delta(t, dist, i, j) for i, j in pairs]
    best = min(deltas) if deltas else 0
    moved = o2.step_best(t, dist, cand)
    assert sorted(t.tolist()) == list(range(n))
    if not moved:
      assert best >= 0
      break
    assert tour_costs(t, dist)[0] == before + best
  else:
    pytest.fail("did not converge")
  if k is None:
    # brute force: no admissible pair improves
    assert all(_spec_delta(t, dist, i, j) >= 0 for i, j in _spec_pairs(n, t, dist, None))


@pytest.mark.parametrize("mode", ["best", "first", "multi_move"])
@pytest.mark.parametrize("k", [None, 4, 20])
def test_c_oracle_matches_numpy_oracle(mode, k):
  n = 70
  dist = rounded_matrix(n, seed=11)
  cand = None if k is None else nearest(dist, k)
  rng = np.random.default_rng(5)
  routes = np.stack([rng.permutation(n) for _ in range(6)]).astype(np.int32)
  for iters in (1, 7, 1000):
    counter = o2.new_counter()
    want = o2.improve_batch(routes, dist, cand, mode, iters, counter)
    got, stats = c_two_opt.improve_batch(routes, dist, cand, mode, iters)
    np.testing.assert_array_equal(got, want)
    assert [stats[1], stats[2]] == [counter.moves, counter.passes]
    if mode != "first":  # first-improvement scans stop at different points (numpy evaluates whole blocks)
      assert stats[0] == counter.deltas
  got, stats = c_two_opt.improve_batch(routes, dist, cand, mode, -1)
  assert stats[1] > 0 and stats[2] >= 6


def test_c_oracle_lattice_ties():
  n = 36
  dist = lattice_matrix(n)
  rng = np.random.default_rng(9)
  routes = np.stack([rng.permutation(n) for _ in range(5)]).astype(np.int32)
  for mode in ("best", "first", "multi_move"):
    want = o2.improve_batch(routes, dist, None, mode, 50)
    got, _ = c_two_opt.improve_batch(routes, dist, None, mode, 50)
    np.testing.assert_array_equal(got, want)


def test_small_and_zero_budget_are_copies():
  dist = rounded_matrix(3, seed=1)
  t = np.array([2, 0, 1], dtype=np.int32)
  out = o2.refine(t, dist, None, "best", 10)
  np.testing.assert_array_equal(out, t)
  assert out is not t
  dist = rounded_matrix(10, seed=1)
  t = np.random.default_rng(1).permutation(10).astype(np.int32)
  np.testing.assert_array_equal(o2.refine(t, dist, None, "best", 0), t)
