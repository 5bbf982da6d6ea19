"""GPU parity of every HIP entry point against the oracle, called through the C ABI.

Bar: bit-exact tours, int32 distances, candidate tables and float64 costs (integer sums).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_two_opt
from oracle import init as oinit
from oracle import operators as ops
from oracle import two_opt as o2
from oracle.distances import candidate_boundary_ties, compute_candidate_list, compute_distance_matrix
from tests.conftest import INT_SPAN, lattice_matrix, nearest, nint_boundary_offsets, rounded_matrix, span_limit_int_coords

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
  print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
  import torch

  from tsp_gpu_clean_amd import _native

  _native.require_device()
  return torch


@pytest.fixture(scope="module")
def nat():
  from tsp_gpu_clean_amd import _native

  return _native


def cu(a):
  import torch

  return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def uniform_coords(n, seed=42, hi=1_000_000):
  return np.random.default_rng(seed).integers(0, hi + 1, (n, 2)).astype(np.float64)


def clustered_coords(n, seed=42, box=1_000_000, clusters=20):
  rng = np.random.default_rng(seed)
  centres = rng.uniform(0, box, (clusters, 2))
  lab = rng.integers(0, clusters, n)
  xy = centres[lab] + rng.normal(0, 0.02 * box, (n, 2))
  return np.rint(np.clip(xy, 0, box))


def device_dist(nat, torch, xy):
  n = xy.shape[0]
  D = torch.empty((n, n), dtype=torch.int32, device="cuda")
  mx = torch.zeros(1, dtype=torch.int32, device="cuda")
  nat.dist_euc2d(cu(xy), D, mx)
  return D, int(mx.item())


# ---- distances and candidate tables -------------------------------------------------------------


def near_half_integer_coords():
  """Points whose distances sit on, or a few ulps / a hair either side of, half-integers (the
  rounding boundary of nint), over magnitudes 1 .. 2^30, plus integer points (t^2, t) whose
  distance from the origin sqrt(t^4 + t^2) lies just below t^2 + 1/2."""
  pts = [(0.0, 0.0)]
  for k in sorted({int(v) for v in np.geomspace(1, 2**30, 60)} | {3, 1000, 999_999, 2**20 + 7}):
    h = k + 0.5
    for d in (0.0, 1e-12, -1e-12, 1e-9, -1e-9, 1e-7, -1e-7, 1e-5, -1e-5):
      pts.append((h + d * h, 0.0))
    lo = hi = h
    for _ in range(3):
      lo, hi = np.nextafter(lo, 0.0), np.nextafter(hi, np.inf)
      pts += [(lo, 0.0), (hi, 0.0)]
  for t in [1, 2, 3, 10, 99, 1000, 3162, 2**15 - 1]:
    pts += [(float(t * t), float(t)), (float(t * t + 1), float(t))]
  return np.array(pts, dtype=np.float64)


@pytest.mark.parametrize("kind", ["uniform1000", "clustered1000", "berlin52", "three_point", "float2100",
                                  "near_half_integers"])
def test_dist_euc2d_matches_oracle(torch_mod, nat, berlin52_coords, kind):
  say(f"dist_euc2d {kind}")
  xy = {
    "uniform1000": lambda: uniform_coords(1000),
    "clustered1000": lambda: clustered_coords(1000),
    "berlin52": lambda: berlin52_coords,
    "three_point": lambda: np.array([[0.0, 0.0], [3.1, 4.0], [1.5, 2.5]]),
    "float2100": lambda: np.random.default_rng(3).random((2100, 2)) * 1000.0,
    "near_half_integers": near_half_integer_coords,
  }[kind]()
  D, mx = device_dist(nat, torch_mod, xy)
  want = compute_distance_matrix(xy)
  np.testing.assert_array_equal(D.cpu().numpy(), want)
  assert mx == int(want.max())


@pytest.mark.parametrize("kind", ["uniform1000", "span_limit", "offset", "boundaries"])
def test_dist_euc2d_i32_matches_oracle(torch_mod, nat, kind):
  """The integer EUC_2D form equals the oracle's FP64 nint entry for entry wherever every pair is
  within the 2^21 span: uniform integers, a box of exactly that span with the nint-boundary pairs,
  large negative coordinates, and the boundary offsets in all quadrants around one centre."""
  say(f"dist_euc2d_i32 {kind}")
  near = nint_boundary_offsets()
  near = near[np.abs(near).max(axis=1) <= INT_SPAN // 2]  # every pair of them within the span
  xy = {
    "uniform1000": lambda: uniform_coords(1000),
    "span_limit": lambda: span_limit_int_coords(1500, seed=1),
    "offset": lambda: uniform_coords(800, seed=4) - float(2**30),
    "boundaries": lambda: np.concatenate([[[0, 0]], near]).astype(np.float64),
  }[kind]()
  n = xy.shape[0]
  D = torch_mod.empty((n, n), dtype=torch_mod.int32, device="cuda")
  bad = torch_mod.zeros(1, dtype=torch_mod.int32, device="cuda")
  nat.dist_euc2d_i32(cu(xy.astype(np.int32)), D, bad)
  np.testing.assert_array_equal(D.cpu().numpy(), compute_distance_matrix(xy))
  assert int(bad.item()) == 0


def test_dist_euc2d_i32_flags_pairs_beyond_the_span(torch_mod, nat):
  say("dist_euc2d_i32 beyond the span")
  xy = np.array([[0, 0], [INT_SPAN, 3], [INT_SPAN + 1, 0], [5, -INT_SPAN - 2]], dtype=np.int32)
  D = torch_mod.empty((4, 4), dtype=torch_mod.int32, device="cuda")
  bad = torch_mod.zeros(1, dtype=torch_mod.int32, device="cuda")
  nat.dist_euc2d_i32(cu(xy), D, bad)
  got, want = D.cpu().numpy(), compute_distance_matrix(xy.astype(np.float64))
  far = np.abs(xy[:, None, :].astype(np.int64) - xy[None, :, :]).max(axis=2) > INT_SPAN
  assert (got[far] == -1).all() and np.array_equal(got[~far], want[~far])
  assert int(bad.item()) == int(far.sum())


@pytest.mark.parametrize("n,k", [(1000, 20), (1000, 32), (52, 5), (7, 20), (2, 3)])
def test_knn_matches_oracle(torch_mod, nat, n, k):
  say(f"knn n={n} k={k}")
  xy = uniform_coords(n, seed=n + k)
  D, _ = device_dist(nat, torch_mod, xy)
  k_eff = min(k, n - 1)
  out = torch_mod.empty((n, k_eff), dtype=torch_mod.int32, device="cuda")
  ties = torch_mod.zeros(1, dtype=torch_mod.int32, device="cuda")
  nat.knn(D, k_eff, out, ties)
  d = compute_distance_matrix(xy)
  want = compute_candidate_list(d, k)
  got = out.cpu().numpy()
  assert int(ties.item()) == candidate_boundary_ties(d, k)
  # every row follows (distance, node index) order ...
  np.testing.assert_array_equal(got, nearest(d, k))
  # ... which is the reference's table on every row without a k-th boundary tie
  dd = d.astype(np.int64)
  np.fill_diagonal(dd, np.iinfo(np.int64).max)
  srt = np.sort(dd, axis=1)
  clean = srt[:, k_eff - 1] != srt[:, k_eff] if k_eff < n - 1 else np.ones(n, bool)
  np.testing.assert_array_equal(got[clean], want[clean])


def test_knn_tie_rows_follow_distance_then_index(torch_mod, nat):
  say("knn lattice ties")
  d = lattice_matrix(49)
  D = cu(d)
  out = torch_mod.empty((49, 6), dtype=torch_mod.int32, device="cuda")
  ties = torch_mod.zeros(1, dtype=torch_mod.int32, device="cuda")
  nat.knn(D, 6, out, ties)
  np.testing.assert_array_equal(out.cpu().numpy(), nearest(d, 6))
  assert int(ties.item()) == candidate_boundary_ties(d, 6)


# ---- tour cost ---------------------------------------------------------------------------------------


def test_tour_cost_known_answers_and_random(torch_mod, nat):
  say("tour_cost")
  d = np.array([[0, 2, 9, 10], [2, 0, 6, 4], [9, 6, 0, 3], [10, 4, 3, 0]], dtype=np.int32)
  t = np.array([[0, 1, 2, 3], [0, 2, 1, 3], [1, 3, 2, 0]], dtype=np.int32)
  out = torch_mod.empty(3, dtype=torch_mod.float64, device="cuda")
  nat.tour_cost(cu(t), cu(d), out)
  assert out.cpu().tolist() == [21.0, 29.0, 18.0]
  xy = uniform_coords(1000)
  dd = compute_distance_matrix(xy)
  rng = np.random.default_rng(1)
  tours = np.stack([rng.permutation(1000) for _ in range(37)]).astype(np.int32)
  out = torch_mod.empty(37, dtype=torch_mod.float64, device="cuda")
  nat.tour_cost(cu(tours), cu(dd), out)
  np.testing.assert_array_equal(out.cpu().numpy(), ops.tour_costs(tours, dd))
  nat.tour_cost(cu(tours), cu(dd), out, closed=False)
  np.testing.assert_array_equal(out.cpu().numpy(), ops.tour_costs(tours, dd, return_to_depot=False))


# ---- 2-opt -------------------------------------------------------------------------------------------


def _run_two_opt(nat, torch, tours, d, cand, mode, iters):
  t = cu(tours.astype(np.int32))
  stats = torch.zeros(3, dtype=torch.int64, device="cuda")
  nat.two_opt(t, cu(d), None if cand is None else cu(cand), {"best": 0, "first": 1, "multi_move": 2}[mode], iters,
              stats)
  return t.cpu().numpy(), stats.cpu().numpy()


@pytest.mark.parametrize("mode", ["best", "first", "multi_move"])
@pytest.mark.parametrize("k", [None, 3, 20])
@pytest.mark.parametrize("n", [4, 5, 6, 8, 10, 17, 40])
def test_two_opt_small_matches_numpy_oracle(torch_mod, nat, mode, k, n):
  say(f"two_opt small mode={mode} k={k} n={n}")
  d = rounded_matrix(n, seed=n)
  cand = None if k is None else nearest(d, k)
  rng = np.random.default_rng(n * 100 + (k or 0))
  tours = np.stack([rng.permutation(n) for _ in range(6)]).astype(np.int32)
  for iters in (1, 3, 10):
    got, _ = _run_two_opt(nat, torch_mod, tours, d, cand, mode, iters)
    np.testing.assert_array_equal(got, o2.improve_batch(tours, d, cand, mode, iters))


@pytest.mark.parametrize("mode", ["best", "first", "multi_move"])
@pytest.mark.parametrize("k", [None, 3])
@pytest.mark.parametrize("n", [9, 16, 25, 64])
def test_two_opt_lattice_ties_match_numpy_oracle(torch_mod, nat, mode, k, n):
  say(f"two_opt lattice mode={mode} k={k} n={n}")
  d = lattice_matrix(n)
  cand = None if k is None else nearest(d, k)
  tours = np.stack([np.random.default_rng(n * 31 + t).permutation(n) for t in range(4)]).astype(np.int32)
  got, _ = _run_two_opt(nat, torch_mod, tours, d, cand, mode, 10)
  np.testing.assert_array_equal(got, o2.improve_batch(tours, d, cand, mode, 10))
  got, _ = _run_two_opt(nat, torch_mod, tours, d, cand, mode, -1)
  want, _ = c_two_opt.improve_batch(tours, d, cand, mode, -1)
  np.testing.assert_array_equal(got, want)


def test_two_opt_duplicate_candidates(torch_mod, nat):
  say("two_opt duplicate candidates")
  n = 30
  d = rounded_matrix(n, seed=5)
  cand = nearest(d, 6)
  cand[:, 3] = cand[:, 1]  # repeated neighbours: pairs appear twice, first occurrence wins ties
  tours = np.stack([np.random.default_rng(t).permutation(n) for t in range(8)]).astype(np.int32)
  for mode in ("best", "first", "multi_move"):
    got, _ = _run_two_opt(nat, torch_mod, tours, d, cand, mode, 20)
    np.testing.assert_array_equal(got, o2.improve_batch(tours, d, cand, mode, 20))


@pytest.mark.parametrize("k", [None, 20])
@pytest.mark.parametrize("mode", ["best", "first", "multi_move"])
def test_two_opt_n1000_to_local_optimum_matches_c_oracle(torch_mod, nat, k, mode):
  say(f"two_opt n=1000 to optimum k={k} mode={mode}")
  xy = uniform_coords(1000)
  d = compute_distance_matrix(xy)
  cand = None if k is None else compute_candidate_list(d, k)
  rows = 6 if k is None else 24
  tours = oinit.random_init(np.arange(1000), rows, oinit.draw_row_seeds(np.random.default_rng(42), rows))
  got, stats = _run_two_opt(nat, torch_mod, tours, d, cand, mode, -1)
  want, wstats = c_two_opt.improve_batch(tours, d, cand, mode, -1)
  np.testing.assert_array_equal(got, want)
  assert stats[1] == wstats[1]  # same number of moves
  if mode != "first":
    assert list(stats) == list(wstats)
  np.testing.assert_array_equal(ops.tour_costs(got, d), ops.tour_costs(want, d))


@pytest.mark.parametrize("n,iters", [(63, -1), (66, -1), (257, -1), (258, 60), (700, 40), (1025, 30), (1026, 30),
                                     (1027, 30), (1283, 20), (1500, 20), (2050, 12)])
def test_two_opt_full_best_slot_phases_match_c_oracle(torch_mod, nat, n, iters):
  """Full-neighbourhood steepest scan around its slot-phase boundaries (n - 2 crossing multiples
  of 64 and 256, both register widths) and on a tie-heavy lattice: tours and counters exact."""
  say(f"two_opt full best phases n={n} iters={iters}")
  d = compute_distance_matrix(uniform_coords(n, seed=n + 3))
  tours = oinit.random_init(np.arange(n), 3, oinit.draw_row_seeds(np.random.default_rng(n), 3))
  got, stats = _run_two_opt(nat, torch_mod, tours, d, None, "best", iters)
  want, wstats = c_two_opt.improve_batch(tours, d, None, "best", iters)
  np.testing.assert_array_equal(got, want)
  assert list(stats) == list(wstats)
  lat = lattice_matrix(n)
  got, stats = _run_two_opt(nat, torch_mod, tours, lat, None, "best", iters)
  want, wstats = c_two_opt.improve_batch(tours, lat, None, "best", iters)
  np.testing.assert_array_equal(got, want)
  assert list(stats) == list(wstats)


def test_two_opt_full_population_is_locally_optimal(torch_mod, nat):
  """P=1024, n=1000, k=20 to local optimum: permutations, no improving candidate pair left."""
  say("two_opt P=1024 n=1000 k=20 to optimum")
  xy = uniform_coords(1000)
  d = compute_distance_matrix(xy)
  cand = compute_candidate_list(d, 20)
  P = 1024
  tours = oinit.random_init(np.arange(1000), P, oinit.draw_row_seeds(np.random.default_rng(42), P))
  got, stats = _run_two_opt(nat, torch_mod, tours, d, cand, "best", -1)
  assert (np.sort(got, axis=1) == np.arange(1000)).all()
  assert (ops.tour_costs(got, d) < ops.tour_costs(tours, d)).all()
  again, _ = c_two_opt.improve_batch(got, d, cand, "best", 1)
  np.testing.assert_array_equal(again, got)
  sample = [0, 511, 1023]
  want, _ = c_two_opt.improve_batch(tours[sample], d, cand, "best", -1)
  np.testing.assert_array_equal(got[sample], want)


def test_two_opt_noop_cases(torch_mod, nat):
  say("two_opt no-op cases")
  d = rounded_matrix(3, seed=1)
  t = np.array([[2, 0, 1]], dtype=np.int32)
  got, stats = _run_two_opt(nat, torch_mod, t, d, None, "best", 10)
  np.testing.assert_array_equal(got, t)
  d = rounded_matrix(12, seed=1)
  t = np.random.default_rng(0).permutation(12).astype(np.int32)[None]
  got, _ = _run_two_opt(nat, torch_mod, t, d, None, "best", 0)
  np.testing.assert_array_equal(got, t)


# ---- GA operators ----------------------------------------------------------------------------------


def _population(n, P, seed):
  rng = np.random.default_rng(seed)
  return np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)


@pytest.mark.parametrize("P,n", [(64, 52), (1024, 1000), (7, 30)])
def test_tournament_pmx_inversion_match_oracle(torch_mod, nat, P, n):
  say(f"ga ops P={P} n={n}")
  pop = _population(n, P, seed=P + n)
  d = compute_distance_matrix(uniform_coords(n, seed=n))
  fit = ops.tour_costs(pop, d)
  fit[::5] = fit[1]  # ties: the first minimum must win
  g = np.random.default_rng(11)
  cand = ops.tournament_draw(g, P, P, 5)
  sel = torch_mod.empty((P, n), dtype=torch_mod.int32, device="cuda")
  win = torch_mod.empty(P, dtype=torch_mod.int32, device="cuda")
  nat.tournament(cu(pop), cu(fit), cu(cand), sel, win)
  np.testing.assert_array_equal(win.cpu().numpy(), ops.tournament_winners(cand, fit))
  np.testing.assert_array_equal(sel.cpu().numpy(), pop[ops.tournament_winners(cand, fit)])
  lo, hi = ops.pmx_draw(g, P, n)
  kids = torch_mod.empty((P, n), dtype=torch_mod.int32, device="cuda")
  nat.pmx(sel, cu(lo), cu(hi), kids)
  want = ops.pmx_apply(pop[ops.tournament_winners(cand, fit)], lo, hi)
  np.testing.assert_array_equal(kids.cpu().numpy(), want)
  probs, p1, p2 = ops.inversion_draw(g, P, n)
  nat.inversion(kids, cu(probs), cu(p1), cu(p2), 0.3)
  np.testing.assert_array_equal(kids.cpu().numpy(), ops.inversion_apply(want, probs, p1, p2, 0.3))


def test_tournaments_follow_argmin_on_nan_fitness(torch_mod, nat):
  """A NaN fitness wins its tournament as in np.argmin (the first NaN), in the tournament kernel
  and in the fused generation's wave tournament."""
  say("tournament NaN fitness")
  P, n = 64, 30
  pop = _population(n, P, seed=9)
  d = compute_distance_matrix(uniform_coords(n, seed=4))
  fit = ops.tour_costs(pop, d)
  fit[[3, 17, 40]] = [np.nan, -np.nan, np.array([0x7FF8000000000005], dtype=np.uint64).view(np.float64)[0]]
  cand = ops.tournament_draw(np.random.default_rng(2), P, P, 5)
  cand[0, 2], cand[1, 4] = 3, 17
  want = ops.tournament_winners(cand, fit)
  sel = torch_mod.empty((P, n), dtype=torch_mod.int32, device="cuda")
  win = torch_mod.empty(P, dtype=torch_mod.int32, device="cuda")
  nat.tournament(cu(pop), cu(fit), cu(cand), sel, win)
  np.testing.assert_array_equal(win.cpu().numpy(), want)
  assert want[0] == 3 and want[1] == 17
  lo, hi = ops.pmx_draw(np.random.default_rng(3), P, n)
  probs, p1, p2 = ops.inversion_draw(np.random.default_rng(4), P, n)
  off = torch_mod.empty((P, n), dtype=torch_mod.int32, device="cuda")
  off_fit = torch_mod.empty(P, dtype=torch_mod.float64, device="cuda")
  nat.generation(cu(pop), cu(fit), cu(d), cu(cand), cu(lo), cu(hi), cu(probs), cu(p1), cu(p2), 0.2, off, off_fit)
  kids = ops.inversion_apply(ops.pmx_apply(pop[want], lo, hi), probs, p1, p2, 0.2)
  np.testing.assert_array_equal(off.cpu().numpy(), kids)


@pytest.mark.parametrize("P,n", [(128, 52), (1024, 1000), (9, 40), (1, 10)])
def test_fused_generation_matches_oracle(torch_mod, nat, P, n):
  say(f"fused generation P={P} n={n}")
  pop = _population(n, P, seed=3 * P + n)
  d = compute_distance_matrix(uniform_coords(n, seed=n + 1))
  fit = ops.tour_costs(pop, d)
  g = np.random.default_rng(5)
  cand = ops.tournament_draw(g, P, P, 5)
  lo, hi = ops.pmx_draw(g, P, n)
  probs, p1, p2 = ops.inversion_draw(g, P, n)
  if P < 2:
    lo, hi = np.zeros(1, np.int32), np.zeros(1, np.int32)
  off = torch_mod.empty((P, n), dtype=torch_mod.int32, device="cuda")
  off_fit = torch_mod.empty(P, dtype=torch_mod.float64, device="cuda")
  nat.generation(cu(pop), cu(fit), cu(d), cu(cand), cu(lo), cu(hi), cu(probs), cu(p1), cu(p2), 0.5, off, off_fit)
  sel = pop[ops.tournament_winners(cand, fit)]
  kids = ops.inversion_apply(ops.pmx_apply(sel, lo, hi) if P >= 2 else sel, probs, p1, p2, 0.5)
  np.testing.assert_array_equal(off.cpu().numpy(), kids)
  np.testing.assert_array_equal(off_fit.cpu().numpy(), ops.tour_costs(kids, d))


def test_rank_select_is_stable_survival(torch_mod, nat):
  say("rank_select survival")
  rng = np.random.default_rng(2)
  old = rng.integers(0, 50, 1024).astype(np.float64)  # many ties
  new = rng.integers(0, 50, 1024).astype(np.float64)
  keys = np.concatenate([old, new])
  order = torch_mod.empty(1024, dtype=torch_mod.int32, device="cuda")
  nat.rank_select(cu(keys), 1024, order)
  zeros = np.zeros_like(keys)
  np.testing.assert_array_equal(order.cpu().numpy(), np.lexsort(np.vstack([keys, zeros, zeros]))[:1024])
  order16 = torch_mod.empty(16, dtype=torch_mod.int32, device="cuda")
  nat.rank_select(cu(old), 16, order16)
  np.testing.assert_array_equal(order16.cpu().numpy(), np.argsort(old, kind="stable")[:16])


@pytest.mark.parametrize("M,keep", [(1, 1), (7, 7), (64, 10), (65, 65), (2048, 1024), (5000, 4999), (20000, 300)])
def test_rank_select_matches_stable_argsort(torch_mod, nat, M, keep):
  """Ties, signed zeros, infinities and NaNs of either sign and several payloads: the order is
  numpy's stable argsort (every NaN last, in index order)."""
  say(f"rank_select M={M} keep={keep}")
  rng = np.random.default_rng(M)
  keys = rng.integers(-20, 20, M).astype(np.float64)
  payload_nan = np.array([0x7FF8000000000001, 0xFFF8000000000002, 0x7FF0000000000003], dtype=np.uint64).view(np.float64)
  special = np.concatenate([[0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1e300, -1e-300], payload_nan])
  hit = rng.random(M) < 0.2
  keys[hit] = special[rng.integers(0, len(special), int(hit.sum()))]
  order = torch_mod.full((keep,), -1, dtype=torch_mod.int32, device="cuda")
  nat.rank_select(cu(keys), keep, order)
  np.testing.assert_array_equal(order.cpu().numpy(), np.argsort(keys, kind="stable")[:keep])


@pytest.mark.parametrize("n", [1, 21, 1000, 1003, 4100])
def test_gather_rows_widths(torch_mod, nat, n):
  """Rows of every width, 16-byte copies (n % 4 == 0) and element copies alike."""
  say(f"gather_rows n={n}")
  A = _population(n, 9, 3)
  B = _population(n, 7, 4)
  src = np.array([15, 2, 8, 0, 9, 3], dtype=np.int32)
  out = torch_mod.empty((6, n), dtype=torch_mod.int32, device="cuda")
  nat.gather_rows(cu(A), cu(np.zeros(9)), cu(B), cu(np.zeros(7)), cu(src), out, None)
  np.testing.assert_array_equal(out.cpu().numpy(), np.vstack([A, B])[src])


@pytest.mark.parametrize("P0,PB,keep,n", [(1024, 1024, 1024, 1000), (16, 16, 16, 52), (7, 5, 9, 21), (3, 0, 2, 8),
                                          (5000, 5000, 4000, 12), (128, 130, 100, 1003)])
def test_survive_equals_rank_select_and_gather(torch_mod, nat, P0, PB, keep, n):
  """The one-launch survival leaves the rows, keys and source order that the stable rank followed
  by the row gather leave (numpy's stable argsort of the keys, ties and NaNs included), with
  16-byte and element copies, and above the small-selection size (M > 8192)."""
  say(f"survive P0={P0} PB={PB} keep={keep} n={n}")
  rng = np.random.default_rng(P0 + PB + n)
  M = P0 + PB
  keys = rng.integers(0, 60, M).astype(np.float64)
  keys[rng.random(M) < 0.05] = np.nan
  A = _population(n, P0, 5)
  B = _population(n, max(PB, 1), 6)[:PB]
  out = torch_mod.full((keep, n), -1, dtype=torch_mod.int32, device="cuda")
  out_fit = torch_mod.empty(keep, dtype=torch_mod.float64, device="cuda")
  order = torch_mod.empty(keep, dtype=torch_mod.int32, device="cuda")
  nat.survive(cu(keys), keep, cu(A), cu(B) if PB else None, out, out_fit, order)
  want = np.argsort(keys, kind="stable")[:keep]
  np.testing.assert_array_equal(order.cpu().numpy(), want)
  np.testing.assert_array_equal(out.cpu().numpy(), np.vstack([A, B])[want] if PB else A[want])
  np.testing.assert_array_equal(out_fit.cpu().numpy(), keys[want])


def test_gather_and_scatter_rows(torch_mod, nat):
  say("gather/scatter rows")
  A = _population(20, 6, 1)
  B = _population(20, 5, 2)
  fa = np.arange(6, dtype=np.float64)
  fb = 10 + np.arange(5, dtype=np.float64)
  src = np.array([7, 0, 10, 3], dtype=np.int32)
  out = torch_mod.empty((4, 20), dtype=torch_mod.int32, device="cuda")
  of = torch_mod.empty(4, dtype=torch_mod.float64, device="cuda")
  nat.gather_rows(cu(A), cu(fa), cu(B), cu(fb), cu(src), out, of)
  both = np.vstack([A, B])
  np.testing.assert_array_equal(out.cpu().numpy(), both[src])
  np.testing.assert_array_equal(of.cpu().numpy(), np.concatenate([fa, fb])[src])
  dst = cu(A.copy())
  dfit = cu(fa.copy())
  nat.scatter_rows(cu(B[:2]), cu(fb[:2]), cu(np.array([4, 1], np.int32)), dst, dfit)
  want = A.copy()
  want[[4, 1]] = B[:2]
  np.testing.assert_array_equal(dst.cpu().numpy(), want)


# ---- initialisers ----------------------------------------------------------------------------------


@pytest.mark.parametrize("P,n", [(1024, 1000), (3, 1), (5, 2)])
def test_random_init_matches_oracle(torch_mod, nat, P, n):
  say(f"random_init P={P} n={n}")
  seeds = oinit.draw_row_seeds(np.random.default_rng(42), P)
  out = torch_mod.empty((P, n), dtype=torch_mod.int32, device="cuda")
  nat.random_init(cu(seeds.view(np.int64)), cu(np.arange(n, dtype=np.int32)), out)
  np.testing.assert_array_equal(out.cpu().numpy(), oinit.random_init(np.arange(n), P, seeds))


@pytest.mark.parametrize("k", [5, 20])
def test_nn_init_matches_oracle(torch_mod, nat, berlin52_coords, k):
  say(f"nn_init k={k}")
  d = compute_distance_matrix(berlin52_coords)
  cand = compute_candidate_list(d, k)  # the kernel consumes the given table
  P, n = 128, 52
  seeds = oinit.draw_row_seeds(np.random.default_rng(42), P)
  out = torch_mod.empty((P, n), dtype=torch_mod.int32, device="cuda")
  work = torch_mod.empty(2 * P * n, dtype=torch_mod.int32, device="cuda")
  nat.nn_init(cu(seeds.view(np.int64)), cu(np.arange(n, dtype=np.int32)), cu(cand), work, out)
  np.testing.assert_array_equal(out.cpu().numpy(), oinit.nn_init(np.arange(n), P, cand, seeds))


def test_nn_init_hand_worked_rows(torch_mod, nat):
  """The two hand-worked GRASP rows of the oracle tests (seed 7, k = 2 and seed 2024, k = 1)."""
  say("nn_init hand-worked rows")
  for seed, cand, want in ((7, [[1, 2], [0, 2], [1, 3], [2, 4], [3, 5], [4, 0]], [2, 1, 0, 4, 3, 5]),
                           (2024, [[1], [0], [0], [0], [0], [0]], [3, 0, 1, 2, 4, 5])):
    out = torch_mod.empty((1, 6), dtype=torch_mod.int32, device="cuda")
    work = torch_mod.empty(12, dtype=torch_mod.int32, device="cuda")
    nat.nn_init(cu(np.array([seed], dtype=np.int64)), cu(np.arange(6, dtype=np.int32)),
                cu(np.array(cand, dtype=np.int32)), work, out)
    assert out.cpu().numpy()[0].tolist() == want


@pytest.mark.parametrize("n,lattice", [(300, False), (400, True), (1, False), (2, False)])
def test_nn_greedy_init_matches_oracle(torch_mod, nat, n, lattice):
  say(f"nn_greedy_init n={n} lattice={lattice}")
  d = lattice_matrix(n) if lattice else compute_distance_matrix(uniform_coords(n, seed=n))
  starts = np.random.default_rng(n).integers(0, n, 6).astype(np.int32)
  out = torch_mod.empty((6, n), dtype=torch_mod.int32, device="cuda")
  nat.nn_greedy_init(cu(starts), cu(d), out)
  np.testing.assert_array_equal(out.cpu().numpy(), oinit.greedy_nn(d, starts))


# ---- packed candidate 2-opt ------------------------------------------------------------------------


def _run_packed(nat, torch, tours, d, cand, iters):
  D = cu(d)
  c = cu(cand)
  table = torch.empty((cand.shape[0], nat.candidate_lanes(cand.shape[1]), 2), dtype=torch.int32, device="cuda")
  nat.candidate_table(c, D, table)
  t = cu(tours.astype(np.int32))
  stats = torch.zeros(3, dtype=torch.int64, device="cuda")
  nat.two_opt_packed(t, D, table, cand.shape[1], iters, stats)
  return t.cpu().numpy(), stats.cpu().numpy(), table.cpu().numpy()


@pytest.mark.parametrize("k", [1, 3, 8, 12, 20, 32])
@pytest.mark.parametrize("n", [5, 17, 40, 300])
def test_packed_candidate_two_opt_matches_oracle(torch_mod, nat, k, n):
  say(f"two_opt_packed n={n} k={k}")
  d = rounded_matrix(n, seed=n + k)
  cand = nearest(d, k)
  tours = np.stack([np.random.default_rng(n * 7 + t).permutation(n) for t in range(6)]).astype(np.int32)
  for iters in (1, 4, 25):
    got, _, table = _run_packed(nat, torch_mod, tours, d, cand, iters)
    np.testing.assert_array_equal(got, o2.improve_batch(tours, d, cand, "best", iters))
  kk = cand.shape[1]
  np.testing.assert_array_equal(table[:, :kk, 0], cand)
  np.testing.assert_array_equal(table[:, :kk, 1], np.take_along_axis(d, cand, axis=1))
  assert (table[:, kk:, 0] == -1).all()


def test_packed_candidate_ties_and_duplicates(torch_mod, nat):
  say("two_opt_packed lattice + duplicates")
  d = lattice_matrix(64)
  cand = nearest(d, 6)
  cand[:, 4] = cand[:, 2]
  tours = np.stack([np.random.default_rng(t).permutation(64) for t in range(8)]).astype(np.int32)
  got, _, _ = _run_packed(nat, torch_mod, tours, d, cand, 30)
  np.testing.assert_array_equal(got, o2.improve_batch(tours, d, cand, "best", 30))


@pytest.mark.parametrize("P", [24, 300, 600])  # 1024-, 512- and 256-thread workgroups on 256 CUs
def test_packed_candidate_n1000_to_optimum(torch_mod, nat, P):
  say(f"two_opt_packed n=1000 P={P} to optimum")
  d = compute_distance_matrix(uniform_coords(1000))
  cand = compute_candidate_list(d, 20)
  tours = oinit.random_init(np.arange(1000), P, oinit.draw_row_seeds(np.random.default_rng(3), P))
  got, stats, _ = _run_packed(nat, torch_mod, tours, d, cand, -1)
  rows = list(range(0, P, max(1, P // 12)))
  want, wstats = c_two_opt.improve_batch(tours[rows], d, cand, "best", -1)
  np.testing.assert_array_equal(got[rows], want)
  same, _ = c_two_opt.improve_batch(got, d, cand, "best", 1)
  np.testing.assert_array_equal(same, got)


@pytest.mark.parametrize("k", [5, 9, 10, 11, 12, 13, 16, 17, 20, 21, 32])
def test_packed_candidate_lane_groups_match_c_oracle(torch_mod, nat, k):
  """Every lanes-per-row group of the many-tour candidate kernel (P > CUs, 512- and 256-thread
  workgroups) and of its coordinate twin: the C oracle's tours and counters."""
  say(f"two_opt_packed lane group k={k}")
  n = 200
  xy = uniform_coords(n, seed=k)
  d = compute_distance_matrix(xy)
  cand = compute_candidate_list(d, k)
  for P, iters in ((300, 6), (600, -1)):
    tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(k + P), P))
    got, stats, _ = _run_packed(nat, torch_mod, tours, d, cand, iters)
    want, wstats = c_two_opt.improve_batch(tours, d, cand, "best", iters)
    np.testing.assert_array_equal(got, want)
    assert stats.tolist() == wstats.tolist()
    D = cu(d)
    table = torch_mod.empty((n, nat.candidate_lanes(k), 2), dtype=torch_mod.int32, device="cuda")
    nat.candidate_table(cu(cand), D, table)
    t = cu(tours)
    nat.two_opt_packed_euc2d(t, cu(xy), table, k, iters)
    np.testing.assert_array_equal(t.cpu().numpy(), want)


@pytest.mark.parametrize("n,P,k,iters", [(1000, 16, 20, 10), (1024, 16, 20, -1), (1025, 16, 20, 10), (300, 300, 8, 5),
                                         (6, 3, 5, 4)])
def test_packed_two_opt_fused_cost(torch_mod, nat, n, P, k, iters):
  """The cost written by the search launch equals the tour cost of its result, for the
  register-resident form (n * k <= 20480, P <= CUs) and the general one alike, matrix and
  coordinates."""
  say(f"two_opt_packed cost n={n} P={P} k={k}")
  xy = uniform_coords(n, seed=n + P)
  D, _ = device_dist(nat, torch_mod, xy)
  cand = torch_mod.empty((n, min(k, n - 1)), dtype=torch_mod.int32, device="cuda")
  nat.knn(D, k, cand)
  kk = cand.shape[1]
  table = torch_mod.empty((n, nat.candidate_lanes(kk), 2), dtype=torch_mod.int32, device="cuda")
  nat.candidate_table(cand, D, table)
  tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(n), P))
  plain, fused, coord = cu(tours), cu(tours), cu(tours)
  s0, s1, s2 = (torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda") for _ in range(3))
  cost = torch_mod.full((P,), -1.0, dtype=torch_mod.float64, device="cuda")
  cost_xy = torch_mod.full((P,), -1.0, dtype=torch_mod.float64, device="cuda")
  nat.two_opt_packed(plain, D, table, kk, iters, s0)
  nat.two_opt_packed(fused, D, table, kk, iters, s1, cost)
  nat.two_opt_packed_euc2d(coord, cu(xy), table, kk, iters, s2, cost_xy)
  np.testing.assert_array_equal(fused.cpu().numpy(), plain.cpu().numpy())
  np.testing.assert_array_equal(coord.cpu().numpy(), plain.cpu().numpy())
  assert s1.cpu().tolist() == s0.cpu().tolist() == s2.cpu().tolist()
  want = torch_mod.empty(P, dtype=torch_mod.float64, device="cuda")
  nat.tour_cost(plain, D, want)
  np.testing.assert_array_equal(cost.cpu().numpy(), want.cpu().numpy())
  np.testing.assert_array_equal(cost_xy.cpu().numpy(), want.cpu().numpy())
  rows = [0, P - 1]
  ref, _ = c_two_opt.improve_batch(tours[rows], compute_distance_matrix(xy), cand.cpu().numpy(), "best", iters)
  np.testing.assert_array_equal(plain.cpu().numpy()[rows], ref)


@pytest.mark.parametrize("n,P,k,iters", [(1000, 16, 20, 10), (1000, 16, 20, -1), (300, 4, 8, -1), (2000, 8, 16, 20),
                                         (6000, 2, 5, 8), (1000, 200, 20, 5), (7, 3, 6, 3)])
@pytest.mark.parametrize("form", ["matrix", "xy"])
def test_two_opt_elite_matches_packed_and_oracle(torch_mod, nat, n, P, k, iters, form):
  """Workgroup clusters per tour (few tours) or the one-workgroup fallback, with the matrix or with
  int32 coordinates (tspmi_two_opt_elite_xy: node records with the successor's coordinates): the
  tours, counters and costs of tspmi_two_opt_packed, the C oracle's tours, and a workspace left
  cleared for the next call (run twice in a row)."""
  say(f"two_opt_elite n={n} P={P} k={k} iters={iters} {form}")
  xy = uniform_coords(n, seed=n * 3 + P)
  D, _ = device_dist(nat, torch_mod, xy)
  kk = min(k, n - 1)
  cand = torch_mod.empty((n, kk), dtype=torch_mod.int32, device="cuda")
  nat.knn(D, kk, cand)
  table = torch_mod.empty((n, nat.candidate_lanes(kk), 2), dtype=torch_mod.int32, device="cuda")
  nat.candidate_table(cand, D, table)
  tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(n + P), P))
  sync = torch_mod.zeros(nat.ELITE_SYNC_WORDS * P, dtype=torch_mod.int64, device="cuda")
  want, got = cu(tours), cu(tours)
  for call in range(2):
    sw = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
    sg = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
    cost = torch_mod.full((P,), -1.0, dtype=torch_mod.float64, device="cuda")
    best = torch_mod.full((1,), -1, dtype=torch_mod.int32, device="cuda")
    nat.two_opt_packed(want, D, table, kk, iters, sw)
    if form == "xy":
      nat.two_opt_elite(got, None, table, kk, iters, sg, cost, sync, best, ixy=cu(xy.astype(np.int32)))
    else:
      nat.two_opt_elite(got, D, table, kk, iters, sg, cost, sync, best)
    torch_mod.cuda.synchronize()
    assert int(best.item()) == int(np.argmin(cost.cpu().numpy()))
    assert not nat.elite_failed(sync)
    words = sync.view(P, nat.ELITE_SYNC_WORDS)
    assert not words[:, 0:256:16].any().item() and not words[:, 1:256:16].any().item(), f"call {call}: slots not cleared"
    assert not words[:, 512:515].any().item(), f"call {call}: error / departure / count words not cleared"
    np.testing.assert_array_equal(got.cpu().numpy(), want.cpu().numpy(), err_msg=f"call {call}")
    assert sg.cpu().tolist() == sw.cpu().tolist()
    ref = torch_mod.empty(P, dtype=torch_mod.float64, device="cuda")
    nat.tour_cost(want, D, ref)
    np.testing.assert_array_equal(cost.cpu().numpy(), ref.cpu().numpy())
  rows = [0, P - 1]
  first, _ = c_two_opt.improve_batch(tours[rows], compute_distance_matrix(xy), cand.cpu().numpy(), "best", iters)
  again, _ = c_two_opt.improve_batch(first, compute_distance_matrix(xy), cand.cpu().numpy(), "best", iters)
  np.testing.assert_array_equal(got.cpu().numpy()[rows], again)


def test_two_opt_elite_error_word_leaves_cluster_untouched(torch_mod, nat):
  """The failure path of the bounded cluster barrier, forced from the host: a tour whose error
  word is set leaves at its first wait (tour and cost untouched, the word reported and kept);
  the other tours are searched as usual, and after the word is cleared the next call refines
  the skipped tour too."""
  say("two_opt_elite error word")
  n, P, k, iters = 600, 8, 12, 6
  xy = uniform_coords(n, seed=77)
  D, _ = device_dist(nat, torch_mod, xy)
  cand = torch_mod.empty((n, k), dtype=torch_mod.int32, device="cuda")
  nat.knn(D, k, cand)
  table = torch_mod.empty((n, nat.candidate_lanes(k), 2), dtype=torch_mod.int32, device="cuda")
  nat.candidate_table(cand, D, table)
  tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(5), P))
  want = cu(tours)
  nat.two_opt_packed(want, D, table, k, iters)
  sync = torch_mod.zeros(nat.ELITE_SYNC_WORDS * P, dtype=torch_mod.int64, device="cuda")
  words = sync.view(P, nat.ELITE_SYNC_WORDS)
  words[3, nat.ELITE_ERROR_WORD] = 1
  got = cu(tours)
  cost = torch_mod.full((P,), -1.0, dtype=torch_mod.float64, device="cuda")
  nat.two_opt_elite(got, D, table, k, iters, None, cost, sync)
  torch_mod.cuda.synchronize()
  assert nat.elite_failed(sync)
  g, w = got.cpu().numpy(), want.cpu().numpy()
  np.testing.assert_array_equal(g[3], tours[3])
  assert cost[3].item() == -1.0
  others = [r for r in range(P) if r != 3]
  np.testing.assert_array_equal(g[others], w[others])
  assert words[:, 0:256:16].sum().item() == 0 and words[:, 1:256:16].sum().item() == 0
  assert words[:, 513].sum().item() == 0
  words[3, nat.ELITE_ERROR_WORD] = 0
  nat.two_opt_elite(got, D, table, k, iters, None, cost, sync)
  torch_mod.cuda.synchronize()
  assert not nat.elite_failed(sync)
  np.testing.assert_array_equal(got.cpu().numpy()[3], w[3])


@pytest.mark.parametrize("n,iters,rows", [(2100, 60, 3), (3000, 40, 2), (6147, 10, 2), (10000, 6, 2), (10243, 4, 1),
                                           (12000, 3, 1)])
def test_two_opt_wide_large_n_matches_c_oracle(torch_mod, nat, n, iters, rows):
  say(f"two_opt wide n={n} iters={iters}")
  d = compute_distance_matrix(uniform_coords(n, seed=n))
  tours = oinit.random_init(np.arange(n), rows, oinit.draw_row_seeds(np.random.default_rng(n), rows))
  got, stats = _run_two_opt(nat, torch_mod, tours, d, None, "best", iters)
  want, wstats = c_two_opt.improve_batch(tours, d, None, "best", iters)
  np.testing.assert_array_equal(got, want)
  assert stats[1] == wstats[1]


def test_two_opt_large_n_lattice_ties_match_c_oracle(torch_mod, nat):
  """Row-staged scan on a tie-heavy lattice (n = 2116): ties resolve to the lowest (i, j)."""
  say("two_opt staged lattice n=2116")
  n = 2116
  d = lattice_matrix(n)
  tours = oinit.random_init(np.arange(n), 2, oinit.draw_row_seeds(np.random.default_rng(5), 2))
  got, stats = _run_two_opt(nat, torch_mod, tours, d, None, "best", 40)
  want, wstats = c_two_opt.improve_batch(tours, d, None, "best", 40)
  np.testing.assert_array_equal(got, want)
  assert list(stats) == list(wstats)


@pytest.fixture(scope="module")
def large_instance():
  n = 25000
  xy = uniform_coords(n, seed=25)
  d = compute_distance_matrix(xy)
  return xy, d, compute_candidate_list(d, 8)


@pytest.mark.parametrize("source", ["coordinates", "matrix"])
def test_two_opt_large_n_matches_c_oracle(torch_mod, nat, large_instance, source):
  """n = 25,000 (beyond the LDS-resident kernels): the workspace kernels' candidate scan (k = 8,
  1,500 passes) and full scan (2 passes) equal the C oracle on one row, tours and counters."""
  import time

  xy, d, cand = large_instance
  n = d.shape[0]
  say(f"two_opt_large n={n} source={source}")
  tours = oinit.random_init(np.arange(n), 3, oinit.draw_row_seeds(np.random.default_rng(25), 3))
  D = cu(d) if source == "matrix" else None
  XY = cu(xy) if source == "coordinates" else None
  tab = torch_mod.empty((n, nat.candidate_lanes(8), 2), dtype=torch_mod.int32, device="cuda")
  nat.candidate_table(cu(cand), D if D is not None else cu(d), tab)
  for table, iters in ((tab, 1500), (None, 2)):
    t = cu(tours)
    stats = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
    cost = torch_mod.empty(3, dtype=torch_mod.float64, device="cuda")
    torch_mod.cuda.synchronize()
    t0 = time.perf_counter()
    nat.two_opt_large(t, D, XY, table, 8 if table is not None else 0, iters, stats, cost)
    torch_mod.cuda.synchronize()
    say(f"  {'k=8' if table is not None else 'full'}: {time.perf_counter() - t0:.3f} s, stats {stats.tolist()}")
    got = t.cpu().numpy()
    want, wstats = c_two_opt.improve_batch(tours[:1], d, cand if table is not None else None, "best", iters)
    np.testing.assert_array_equal(got[:1], want)
    assert (np.sort(got, axis=1) == np.arange(n)).all()
    np.testing.assert_array_equal(cost.cpu().numpy(), ops.tour_costs(got, d))
    if table is None:
      assert int(stats[2].item()) == 3 * iters and int(stats[1].item()) == 3 * iters


# ---- EUC_2D coordinate-mode 2-opt ---------------------------------------------------------------


def lattice_coords(n):
  side = int(np.ceil(np.sqrt(n)))
  return np.array([[x, y] for x in range(side) for y in range(side)][:n], dtype=np.float64)


@pytest.mark.parametrize("n,P,iters,kind", [(5, 6, -1, "uniform"), (40, 8, -1, "float"), (300, 16, -1, "lattice"),
                                            (1000, 64, -1, "uniform"), (1000, 300, -1, "float"),
                                            (2100, 300, 15, "uniform"), (5000, 4, 6, "float"), (8194, 2, 3, "uniform")])
def test_two_opt_euc2d_equals_matrix_kernel(torch_mod, nat, n, P, iters, kind):
  """Distances computed from coordinates give the matrix kernel's tours and counters exactly."""
  say(f"two_opt_euc2d n={n} P={P} iters={iters} {kind}")
  xy = {"uniform": lambda: uniform_coords(n, seed=n), "float": lambda: np.random.default_rng(n).random((n, 2)) * 1000.0,
        "lattice": lambda: lattice_coords(n)}[kind]()
  d = compute_distance_matrix(xy)
  tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(n + 1), P))
  want, wstats = _run_two_opt(nat, torch_mod, tours, d, None, "best", iters)
  t = cu(tours)
  stats = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
  nat.two_opt_euc2d(t, cu(xy), iters, stats)
  np.testing.assert_array_equal(t.cpu().numpy(), want)
  assert stats.cpu().tolist() == list(wstats)
  if n <= 1000:
    ref, _ = c_two_opt.improve_batch(tours[:2], d, None, "best", iters)
    np.testing.assert_array_equal(want[:2], ref)


@pytest.mark.parametrize("n,P,k,iters", [(40, 6, 8, -1), (1000, 24, 20, -1), (1000, 300, 20, -1), (1000, 600, 32, -1),
                                         (10000, 4, 8, 40)])
@pytest.mark.parametrize("form", ["f64", "i32", "i32_span_limit", "i32_lattice"])
def test_two_opt_packed_euc2d_equals_matrix_kernel(torch_mod, nat, n, P, k, iters, form):
  """Candidate scan with every distance from coordinates, float64 (FP64 form) or int32 (integer
  form: uniform integers, a box of exactly the 2^20 span with nint-boundary pairs, a tie-heavy
  lattice): tours and counters equal the matrix kernel's."""
  say(f"two_opt_packed_euc2d n={n} P={P} k={k} {form}")
  xy = {"f64": lambda: uniform_coords(n, seed=3 * n), "i32": lambda: uniform_coords(n, seed=3 * n),
        "i32_span_limit": lambda: span_limit_int_coords(n, seed=n),
        "i32_lattice": lambda: lattice_coords(n)}[form]()
  D, _ = device_dist(nat, torch_mod, xy)
  X = cu(xy.astype(np.int32)) if form.startswith("i32") else cu(xy)
  cand = torch_mod.empty((n, k), dtype=torch_mod.int32, device="cuda")
  nat.knn(D, k, cand)
  table = torch_mod.empty((n, nat.candidate_lanes(k), 2), dtype=torch_mod.int32, device="cuda")
  nat.candidate_table(cand, D, table)
  tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(n + 2), P))
  a, b = cu(tours), cu(tours)
  sa = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
  sb = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
  nat.two_opt_packed(a, D, table, k, iters, sa)
  nat.two_opt_packed_euc2d(b, X, table, k, iters, sb)
  np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy())
  assert sb.cpu().tolist() == sa.cpu().tolist()


@pytest.mark.parametrize("n,P,iters", [(4, 3, -1), (5, 6, -1), (10, 6, -1), (63, 4, -1), (66, 4, -1), (257, 3, -1),
                                       (258, 3, 60), (700, 3, 40), (1025, 3, 30), (1026, 3, 30), (1283, 2, 20),
                                       (2050, 2, 12)])
@pytest.mark.parametrize("kind", ["uniform", "float", "lattice", "span_limit", "beyond_span", "stacked"])
def test_two_opt_full_euc2d_matches_c_oracle(torch_mod, nat, n, P, iters, kind):
  """Full-neighbourhood steepest scan with rows from the coordinates, around the slot-phase and
  16-row mix-period boundaries, on integer (lower-bound rows, exact re-walk of the candidate
  diagonals), fractional (the FP64 form) and tie-heavy lattice instances, on integers spanning
  exactly the integer form's limit with nint-boundary pairs, on integers one past that span (FP64
  form again), and on points stacked on 12 sites (most deltas tie, so many lanes are candidates
  every pass): tours and counters equal the C oracle's on the matrix of the same coordinates."""
  say(f"two_opt_full_euc2d n={n} P={P} iters={iters} {kind}")

  def beyond():
    xy = uniform_coords(n, seed=n + 13, hi=INT_SPAN + 1)
    xy[0], xy[-1] = (0.0, 0.0), (float(INT_SPAN + 1), float(INT_SPAN + 1))
    return xy

  xy = {"uniform": lambda: uniform_coords(n, seed=n + 11),
        "float": lambda: np.random.default_rng(n).random((n, 2)) * 1000.0,
        "lattice": lambda: lattice_coords(n),
        "span_limit": lambda: span_limit_int_coords(n, seed=n),
        "beyond_span": beyond,
        "stacked": lambda: uniform_coords(12, seed=n + 17)[np.random.default_rng(n).integers(0, 12, n)]}[kind]()
  d = compute_distance_matrix(xy)
  D, _ = device_dist(nat, torch_mod, xy)
  np.testing.assert_array_equal(D.cpu().numpy(), d)
  tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(n + 5), P))
  t = cu(tours)
  stats = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
  nat.two_opt_full_euc2d(t, D, cu(xy), iters, stats)
  want, wstats = c_two_opt.improve_batch(tours, d, None, "best", iters)
  np.testing.assert_array_equal(t.cpu().numpy(), want)
  assert stats.cpu().tolist() == list(wstats)


@pytest.mark.parametrize("n,P,iters,kind", [(4, 3, -1, "uniform"), (10, 6, -1, "uniform"), (257, 3, -1, "lattice"),
                                             (700, 3, 40, "float"), (1026, 3, 30, "stacked"), (2050, 2, 12, "span_limit"),
                                             (1000, 64, -1, "uniform")])
def test_two_opt_full_euc2d_ordered_equals_plain(torch_mod, nat, n, P, iters, kind):
  """The Hilbert-renumbered scan (tspmi_two_opt_full_euc2d_ordered): the same tours and counters as
  the scan on the instance as given, on integer, fractional, tie-heavy and span-limit instances;
  a workspace one byte short is refused before the tours are touched."""
  say(f"two_opt_full_euc2d_ordered n={n} P={P} iters={iters} {kind}")
  xy = {"uniform": lambda: uniform_coords(n, seed=n + 21),
        "float": lambda: np.random.default_rng(n).random((n, 2)) * 1000.0,
        "lattice": lambda: lattice_coords(n),
        "span_limit": lambda: span_limit_int_coords(n, seed=n + 1),
        "stacked": lambda: uniform_coords(12, seed=n)[np.random.default_rng(n + 3).integers(0, 12, n)]}[kind]()
  D, _ = device_dist(nat, torch_mod, xy)
  tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(n + 6), P))
  a, b = cu(tours), cu(tours)
  sa = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
  sb = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
  nat.two_opt_full_euc2d(a, D, cu(xy), iters, sa)
  ws = torch_mod.empty(nat.full_euc2d_ws_bytes(n), dtype=torch_mod.uint8, device="cuda")
  nat.two_opt_full_euc2d(b, D, cu(xy), iters, sb, ws=ws)
  np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy())
  assert sb.cpu().tolist() == sa.cpu().tolist()
  c = cu(tours)
  with pytest.raises(ValueError, match="workspace"):
    nat.two_opt_full_euc2d(c, D, cu(xy), iters, None, ws=ws[: ws.numel() - 1])
  np.testing.assert_array_equal(c.cpu().numpy(), tours)


@pytest.mark.parametrize("n,P,k,iters,kind", [(40, 6, 8, -1, "uniform"), (1000, 24, 20, -1, "lattice"),
                                               (1000, 300, 20, 30, "uniform"), (10000, 4, 8, 40, "span_limit")])
def test_two_opt_packed_euc2d_ordered_equals_plain(torch_mod, nat, n, P, k, iters, kind):
  """The Hilbert-renumbered candidate scan (tspmi_two_opt_packed_euc2d_i32_ordered): the tours,
  counters, costs and per-move cost log of the scan on the instance as given."""
  say(f"two_opt_packed_euc2d ordered n={n} P={P} k={k} {kind}")
  xy = {"uniform": lambda: uniform_coords(n, seed=n + 31), "lattice": lambda: lattice_coords(n),
        "span_limit": lambda: span_limit_int_coords(n, seed=n + 2)}[kind]()
  D, _ = device_dist(nat, torch_mod, xy)
  cand = torch_mod.empty((n, k), dtype=torch_mod.int32, device="cuda")
  nat.knn(D, k, cand)
  table = torch_mod.empty((n, nat.candidate_lanes(k), 2), dtype=torch_mod.int32, device="cuda")
  nat.candidate_table(cand, D, table)
  ixy = cu(xy.astype(np.int32))
  tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(n + 9), P))
  ws = torch_mod.empty(nat.packed_euc2d_ws_bytes(n, k), dtype=torch_mod.uint8, device="cuda")
  out = {}
  for form in ("plain", "ordered"):
    t = cu(tours)
    st = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
    cost = torch_mod.zeros(P, dtype=torch_mod.float64, device="cuda")
    log = torch_mod.zeros((P, 6, 2), dtype=torch_mod.int64, device="cuda")
    nat.two_opt_packed_euc2d_log(t, ixy, table, k, iters, log, st, cost, ws=ws if form == "ordered" else None)
    out[form] = (t.cpu().numpy(), st.cpu().tolist(), cost.cpu().numpy(), log[:, :, 0].cpu().numpy())
  for a, b in zip(out["plain"], out["ordered"]):
    np.testing.assert_array_equal(np.asarray(b), np.asarray(a))


def test_two_opt_full_euc2d_population_equals_matrix_kernel(torch_mod, nat):
  """P=1024, n=1000 to the local optimum: the same tours and counters as the matrix-only scan."""
  say("two_opt_full_euc2d P=1024 n=1000 to optimum")
  n, P = 1000, 1024
  xy = uniform_coords(n, seed=7)
  D, _ = device_dist(nat, torch_mod, xy)
  tours = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(np.random.default_rng(8), P))
  a, b = cu(tours), cu(tours)
  sa = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
  sb = torch_mod.zeros(3, dtype=torch_mod.int64, device="cuda")
  nat.two_opt(a, D, None, 0, -1, sa)
  nat.two_opt_full_euc2d(b, D, cu(xy), -1, sb)
  np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy())
  assert sb.cpu().tolist() == sa.cpu().tolist()
  sample = [0, 1023]
  want, _ = c_two_opt.improve_batch(tours[sample], D.cpu().numpy(), None, "best", -1)
  np.testing.assert_array_equal(b.cpu().numpy()[sample], want)


def test_two_opt_full_euc2d_rejects_bad_input(torch_mod, nat):
  xy = uniform_coords(2051)
  D, _ = device_dist(nat, torch_mod, xy)
  with pytest.raises(ValueError):  # above the diagonal scan's 8 slots
    nat.two_opt_full_euc2d(cu(np.arange(2051, dtype=np.int32)[None, :]), D, cu(xy), 1)
  with pytest.raises(ValueError):  # coordinates of another instance size
    nat.two_opt_full_euc2d(cu(np.arange(10, dtype=np.int32)[None, :]), D[:10, :10].contiguous(), cu(xy[:11]), 1)
  t = cu(np.arange(3, dtype=np.int32)[None, :])
  nat.two_opt_full_euc2d(t, D[:3, :3].contiguous(), cu(xy[:3]), 5)  # n < 4: the tour is kept
  assert t.cpu().tolist() == [[0, 1, 2]]


def test_two_opt_euc2d_rejects_bad_input(torch_mod, nat):
  t = cu(np.arange(9000, dtype=np.int32)[None, :])
  with pytest.raises(ValueError):
    nat.two_opt_euc2d(t, cu(uniform_coords(9000)), 1)  # above the full-neighbourhood size limit
  with pytest.raises(ValueError):
    nat.two_opt_euc2d(cu(np.arange(10, dtype=np.int32)[None, :]), cu(uniform_coords(11)), 1)


# ---- order crossover and swap mutation ------------------------------------------------------------


@pytest.mark.parametrize("P,n", [(1, 50), (2, 2), (7, 3), (64, 50), (1024, 1000), (6, 5000)])
def test_ox_matches_oracle(torch_mod, nat, P, n):
  say(f"ox P={P} n={n}")
  rng = np.random.default_rng(P * 7 + n)
  pop = np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)
  lo, hi = ops.pmx_draw(rng, P, n) if P >= 2 else (np.zeros(1, np.int32), np.zeros(1, np.int32))
  if P >= 6:  # extreme cuts: whole tour, single gene at both ends
    lo[:3] = [0, 0, n - 1]
    hi[:3] = [n - 1, 0, n - 1]
  out = torch_mod.empty((P, n), dtype=torch_mod.int32, device="cuda")
  nat.ox(cu(pop), cu(lo), cu(hi), out)
  np.testing.assert_array_equal(out.cpu().numpy(), ops.ox_apply(pop, lo, hi))


@pytest.mark.parametrize("P,n", [(1, 2), (100, 50), (1024, 1000)])
def test_swap_matches_oracle(torch_mod, nat, P, n):
  say(f"swap P={P} n={n}")
  rng = np.random.default_rng(P + n)
  pop = np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)
  probs, p1, p2 = ops.inversion_draw(rng, P, n)
  t = cu(pop)
  nat.swap(t, cu(probs), cu(p1), cu(p2), 0.3)
  np.testing.assert_array_equal(t.cpu().numpy(), ops.swap_apply(pop, probs, p1, p2, 0.3))


# ---- TSPLIB metrics ------------------------------------------------------------------------------


@pytest.mark.parametrize("metric", ["EUC_2D", "CEIL_2D", "MAN_2D", "MAX_2D", "ATT", "GEO"])
def test_dist_metric_matches_oracle(torch_mod, nat, metric):
  say(f"dist_metric {metric}")
  rng = np.random.default_rng(11)
  if metric == "GEO":  # DDD.MM latitudes / longitudes
    deg = np.stack([rng.integers(-89, 89, 700), rng.integers(-179, 179, 700)], axis=1).astype(np.float64)
    xy = deg + np.sign(deg + 0.5) * rng.integers(0, 60, (700, 2)) / 100.0
  else:
    xy = np.concatenate([rng.random((400, 2)) * 1000.0, rng.integers(-5000, 5000, (300, 2)).astype(np.float64)])
  want = compute_distance_matrix(xy, metric)
  D = torch_mod.empty((xy.shape[0], xy.shape[0]), dtype=torch_mod.int32, device="cuda")
  mx = torch_mod.zeros(1, dtype=torch_mod.int32, device="cuda")
  nat.dist_metric(cu(xy), metric, D, mx)
  np.testing.assert_array_equal(D.cpu().numpy(), want)
  assert int(mx.item()) == int(want.max())
  pts = np.array([[16.47, 96.10], [16.47, 94.44], [20.09, 92.54]]) if metric == "GEO" else np.array(
    [[0.0, 0.0], [3.1, 4.0], [1.5, 2.5]])
  small = torch_mod.empty((3, 3), dtype=torch_mod.int32, device="cuda")
  nat.dist_metric(cu(pts), metric, small)
  np.testing.assert_array_equal(small.cpu().numpy(), compute_distance_matrix(pts, metric))


@pytest.mark.parametrize("mode", ["best", "first", "multi_move"])
@pytest.mark.parametrize("k", [None, 6])
def test_two_opt_asymmetric_matrix_matches_oracle(torch_mod, nat, mode, k):
  """Explicit asymmetric (ATSP-style) matrices: the same delta formula and scan as the reference.

  Bounded pass counts only: the symmetric delta formula does not make the tour cost decrease on
  an asymmetric matrix, so improving moves can cycle and "to the local optimum" need not end
  (the C oracle still finds a move on every one of 20000 passes for these inputs)."""
  say(f"two_opt asymmetric mode={mode} k={k}")
  n = 90
  rng = np.random.default_rng(17)
  d = rng.integers(1, 10_000, (n, n)).astype(np.int32)
  np.fill_diagonal(d, 0)
  cand = None if k is None else nearest(d, k)
  tours = np.stack([rng.permutation(n) for _ in range(8)]).astype(np.int32)
  for iters in (3, 25):
    got, _ = _run_two_opt(nat, torch_mod, tours, d, cand, mode, iters)
    np.testing.assert_array_equal(got, o2.improve_batch(tours, d, cand, mode, iters))
  out = torch_mod.empty(8, dtype=torch_mod.float64, device="cuda")
  nat.tour_cost(cu(tours), cu(d), out)
  np.testing.assert_array_equal(out.cpu().numpy(), ops.tour_costs(tours, d))
