"""GPU parity of the learned-scorer kernels against oracle/mlp.py, called through the C ABI.

Bars: bf16 conversion, features (as bf16 bits), labels and top-k bit-exact; the MLP forward
within 1e-3 * (1 + |z|) per sample (fp32 accumulation against float64 at the same bf16 rounding
points; a rare one-ulp bf16 flip of an activation is absorbed by that tolerance); the gradient
within relative L2 error 1e-2 per parameter block; Adam within 1e-5 relative (float32 both sides,
the device fusing multiply-adds).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp
from oracle.distances import compute_candidate_list, compute_distance_matrix

pytestmark = pytest.mark.gpu

BLOCKS = ((mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_B2, mlp.OFF_W3),
          (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.NPARAM))


def say(msg: str) -> None:
  print(msg, flush=True)


@pytest.fixture(scope="module")
def env():
  import torch

  from tsp_gpu_clean_amd import _native

  _native.require_device()
  return torch, _native


def _theta(seed: int) -> np.ndarray:
  th = mlp.init_params(np.random.default_rng(seed))
  rng = np.random.default_rng(seed + 100)
  th[mlp.OFF_B1 : mlp.OFF_W2] = rng.normal(0, 0.1, mlp.H)
  th[mlp.OFF_B2 : mlp.OFF_W3] = rng.normal(0, 0.1, mlp.H)
  th[mlp.OFF_B3] = -0.3
  return th.astype(np.float32)


def _inputs(S: int, seed: int):
  rng = np.random.default_rng(seed)
  x = np.zeros((S, mlp.F), np.float32)
  x[:, : mlp.NFEAT] = rng.normal(0, 1, (S, mlp.NFEAT))
  y = (rng.random(S) < 0.25).astype(np.float32)
  return mlp.bf16_bits(x), y


def _dev16(torch, bits):
  return torch.from_numpy(np.ascontiguousarray(bits).view(np.int16)).cuda()


def test_param_count_and_to_bf16(env):
  torch, nat = env
  assert nat.scorer_param_count() == mlp.NPARAM
  rng = np.random.default_rng(1)
  v = np.concatenate([rng.normal(0, 3, 5000), [1 + 2**-8, 1 + 3 * 2**-8, -0.0, 0.0, 1e-38]]).astype(np.float32)
  out = torch.empty(v.size, dtype=torch.int16, device="cuda")
  nat.to_bf16(torch.from_numpy(v).cuda(), out)
  assert np.array_equal(out.cpu().numpy().view(np.uint16), mlp.bf16_bits(v))


@pytest.mark.parametrize("S", [32, 1000, 4096 + 40])
def test_forward_matches_oracle(env, S):
  torch, nat = env
  say(f"mlp forward S={S}")
  xb, _ = _inputs(S, S)
  wb = mlp.bf16_bits(_theta(3))
  z = torch.full((S,), float("nan"), dtype=torch.float32, device="cuda")
  nat.mlp_forward(_dev16(torch, xb.reshape(S, mlp.F)), _dev16(torch, wb), z)
  got = z.cpu().numpy().astype(np.float64)
  want, _ = mlp.forward(xb, wb)
  err = np.abs(got - want)
  assert np.isfinite(got).all()
  assert (err <= 1e-3 * (1 + np.abs(want))).all(), float(err.max())
  assert np.median(err) <= 1e-5 * (1 + np.median(np.abs(want)))


@pytest.mark.parametrize("S", [128, 2048 + 256, 3 * 2048 + 8])
def test_gradient_matches_oracle(env, S):
  torch, nat = env
  say(f"mlp grad S={S}")
  xb, y = _inputs(S, 7 + S)
  wb = mlp.bf16_bits(_theta(5))
  w, p = nat.mlp_work_sizes(S)
  work = torch.empty(w, dtype=torch.int16, device="cuda")
  partials = torch.empty(p, dtype=torch.float32, device="cuda")
  grad = torch.full((mlp.NPARAM + 1,), float("nan"), dtype=torch.float32, device="cuda")
  nat.mlp_grad(_dev16(torch, xb.reshape(S, mlp.F)), torch.from_numpy(y).cuda(), _dev16(torch, wb), work, partials, grad)
  g = grad.cpu().numpy().astype(np.float64)
  loss, want = mlp.loss_and_grad(xb, y, wb)
  assert g[mlp.NPARAM] == pytest.approx(loss, rel=1e-4)
  for lo, hi in BLOCKS:
    e = np.linalg.norm(g[lo:hi] - want[lo:hi]) / max(np.linalg.norm(want[lo:hi]), 1e-30)
    assert e < 1e-2, (lo, hi, e)
  pad = (np.arange(mlp.H * mlp.F) % mlp.F) >= mlp.NFEAT
  assert (g[mlp.OFF_W1 : mlp.OFF_B1][pad] == 0).all()


def test_gradient_is_deterministic(env):
  torch, nat = env
  S = 4096
  xb, y = _inputs(S, 31)
  wb = _dev16(torch, mlp.bf16_bits(_theta(6)))
  X = _dev16(torch, xb.reshape(S, mlp.F))
  yy = torch.from_numpy(y).cuda()
  w, p = nat.mlp_work_sizes(S)
  work = torch.empty(w, dtype=torch.int16, device="cuda")
  partials = torch.empty(p, dtype=torch.float32, device="cuda")
  outs = []
  for _ in range(2):
    grad = torch.empty(mlp.NPARAM + 1, dtype=torch.float32, device="cuda")
    nat.mlp_grad(X, yy, wb, work, partials, grad)
    outs.append(grad.cpu().numpy())
  assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_adam_matches_oracle(env):
  torch, nat = env
  rng = np.random.default_rng(2)
  th = _theta(8)
  m = rng.normal(0, 0.01, mlp.NPARAM).astype(np.float32)
  v = np.abs(rng.normal(0, 0.001, mlp.NPARAM)).astype(np.float32)
  g = rng.normal(0, 0.1, mlp.NPARAM).astype(np.float32)
  dth, dm, dv = (torch.from_numpy(a.copy()).cuda() for a in (th, m, v))
  grad = torch.from_numpy(np.concatenate([g, [0.0]]).astype(np.float32)).cuda()
  wbf = torch.empty(mlp.NPARAM, dtype=torch.int16, device="cuda")
  nat.adam(dth, dm, dv, grad, 3, 1e-3, 0.9, 0.999, 1e-8, 1e-4, wbf)
  th2, m2, v2 = mlp.adam(th, m, v, g, 3, lr=1e-3, wd=1e-4)
  # float32 both sides; the device may fuse a multiply-add (one rounding fewer)
  assert np.allclose(dth.cpu().numpy(), th2, rtol=1e-6, atol=1e-7)
  assert np.allclose(dm.cpu().numpy(), m2, rtol=1e-5, atol=1e-8)
  assert np.allclose(dv.cpu().numpy(), v2, rtol=1e-5, atol=1e-11)
  assert np.array_equal(wbf.cpu().numpy().view(np.uint16), mlp.bf16_bits(dth.cpu().numpy()))
  pad = (np.arange(mlp.H * mlp.F) % mlp.F) >= mlp.NFEAT
  assert (dth.cpu().numpy()[mlp.OFF_W1 : mlp.OFF_B1][pad] == 0).all()


def _instance(n: int, seed: int, clustered: bool):
  rng = np.random.default_rng(seed)
  if clustered:
    c = rng.uniform(0, 1e5, (6, 2))
    xy = np.rint(np.clip(c[rng.integers(0, 6, n)] + rng.normal(0, 3e3, (n, 2)), 0, 1e5))
  else:
    xy = rng.integers(0, 100_001, (n, 2)).astype(np.float64)
  D = compute_distance_matrix(xy)
  return xy, D, compute_candidate_list(D, 32)


@pytest.mark.parametrize("n,clustered", [(40, False), (300, True), (777, False)])
def test_features_match_oracle(env, n, clustered):
  torch, nat = env
  say(f"scorer features n={n}")
  xy, D, nn = _instance(n, n, clustered)
  X = torch.empty((n * 32, 16), dtype=torch.int16, device="cuda")
  nat.scorer_features(torch.from_numpy(xy).cuda(), torch.from_numpy(nn.astype(np.int32)).cuda(),
                      torch.from_numpy(D).cuda(), X)
  got = X.cpu().numpy().view(np.uint16)
  want = mlp.bf16_bits(mlp.features(xy, nn, D))
  assert np.array_equal(got, want)


def test_labels_and_topk(env):
  torch, nat = env
  n = 500
  xy, D, nn = _instance(n, 4, False)
  rng = np.random.default_rng(4)
  # greedy nearest-neighbour walk, so that many candidate pairs are tour edges
  seen = np.zeros(n, bool)
  tour = [0]
  seen[0] = True
  for _ in range(n - 1):
    row = np.where(seen, np.iinfo(np.int32).max, D[tour[-1]])
    tour.append(int(np.argmin(row)))
    seen[tour[-1]] = True
  tour = np.array(tour, np.int32)
  y = torch.full((n * 32,), -1.0, dtype=torch.float32, device="cuda")
  work = torch.empty(n, dtype=torch.int32, device="cuda")
  nn_d = torch.from_numpy(nn.astype(np.int32)).cuda()
  nat.scorer_labels(torch.from_numpy(tour).cuda(), nn_d, work, y)
  succ = np.empty(n, np.int64)
  succ[tour] = np.roll(tour, -1)
  u = np.repeat(np.arange(n), 32)
  v = nn.reshape(-1)
  want = ((succ[u] == v) | (succ[v] == u)).astype(np.float32)
  assert np.array_equal(y.cpu().numpy(), want)
  # top-k: scores with ties; higher first, lower rank on ties, NaN never chosen
  z = np.round(rng.normal(0, 1, n * 32), 1).astype(np.float32)
  z[::97] = np.nan
  for k in (1, 8, 20):
    out = torch.empty((n, k), dtype=torch.int32, device="cuda")
    nat.scorer_topk(torch.from_numpy(z).cuda(), nn_d, k, out)
    zz = z.reshape(n, 32).astype(np.float64)
    key = np.where(np.isnan(zz), np.inf, -zz)
    order = np.lexsort((np.broadcast_to(np.arange(32), (n, 32)), key), axis=1)[:, :k]
    ref = np.take_along_axis(nn, order, axis=1)
    ref = np.where(np.take_along_axis(np.isnan(zz), order, axis=1), -1, ref)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_training_reduces_loss(env):
  """A few hundred device Adam steps on labelled 2-opt tours lower the loss and beat the prior."""
  torch, nat = env
  from tsp_gpu_clean_amd.scorer import LearnedCandidateScorer, build_training_set, training_instances

  inst = training_instances(count=8, seed=3, lo=200, hi=600)
  ts = build_training_set(inst, population=8, generations=2, chunk=8)
  say(f"training set: {ts.X.shape[0]} samples, {ts.positives} positives, {ts.seconds:.2f}s")
  assert ts.positives > 0
  sc = LearnedCandidateScorer(seed=0)
  losses = sc.train(ts.X, ts.y, steps=200, batch=8192, lr=3e-3, log_every=20)
  prior = ts.positives / ts.X.shape[0]
  base = -(prior * np.log(prior) + (1 - prior) * np.log(1 - prior))
  assert losses[-1] < losses[0]
  assert losses[-1] < base


def test_learned_candidates_gain_over_knn_at_equal_width(env):
  """What the trained network gains: with the same 8 candidates per city, 2-opt on the learned
  lists reaches better local optima than on the 8 nearest neighbours (mean over 32 start tours,
  same starts), measured against full-neighbourhood 2-opt from those starts."""
  torch, nat = env
  from tsp_gpu_clean_amd.scorer import LearnedCandidateScorer, build_training_set, evaluate_gap, training_instances

  inst = training_instances(count=24, seed=21, lo=300, hi=1200)
  ts = build_training_set(inst, population=8, generations=2, chunk=24)
  sc = LearnedCandidateScorer(seed=1)
  losses = sc.train(ts.X, ts.y, steps=400, batch=16384, lr=3e-3, log_every=100)
  say(f"gain test: {ts.X.shape[0]} samples, losses {losses}")
  res = evaluate_gap(n=2000, P=32, scorer=sc, k_learned=8, seed=77, full_chunk=100, cand_chunk=50)
  say(f"gain test result: {res}")
  full = res["full"]["mean"]
  learned = res["methods"]["learned8"]["mean"]
  knn8 = res["methods"]["knn8"]["mean"]
  assert learned < knn8, (learned, knn8, full)
  assert res["methods"]["learned8"]["best"] <= 1.05 * res["full"]["best"]


def test_local_search_strategy_with_learned_candidates(env):
  """The drop-in strategy with the shipped scorer: bit-exact against the oracle 2-opt run on the
  same learned lists, in best and multi_move modes."""
  torch, nat = env
  import tsp_gpu_clean_amd as t
  from oracle import two_opt as o2
  from tsp_gpu_clean_amd.scorer import LearnedCandidateScorer

  rng = np.random.default_rng(12)
  xy = rng.integers(0, 1_000_001, (600, 2)).astype(np.float64)
  ctx = t.ProblemContext(t.Problem.from_coordinates("u600", xy), seed=1)
  sc = LearnedCandidateScorer.load()
  routes = np.stack([rng.permutation(600) for _ in range(8)]).astype(np.int32)
  for mode in ("best", "multi_move"):
    ls = t.LocalSearchGPUTourTwoOpt(max_iterations=25, mode=mode, candidate_k=8, candidate_scorer=sc)
    got = ls.improve_batch(routes, ctx)
    cand = ctx.get_gpu_learned_candidates(sc, 8)[0].cpu().numpy()
    want = o2.improve_batch(routes, ctx.get_cpu_distances(), cand, mode, 25)
    np.testing.assert_array_equal(got, want)
  knn = ctx.get_cpu_candidate_list(8)
  assert not np.array_equal(np.sort(cand, axis=1), np.sort(knn, axis=1))  # the lists really differ


def test_pass_log_integer_form_equals_fp64_form(env):
  """The logged candidate scan from int32 coordinates (integer distance form) leaves the same
  tours, counters and final costs and logs the same cost after every move as from float64."""
  torch, nat = env
  n, P, k = 3000, 24, 8
  xy_h = np.random.default_rng(6).integers(0, 1_000_001, (n, 2)).astype(np.float64)
  xy = torch.from_numpy(xy_h).cuda()
  ixy = torch.from_numpy(xy_h.astype(np.int32)).cuda()
  D = torch.empty((n, n), dtype=torch.int32, device="cuda")
  nat.dist_euc2d(xy, D)
  cand = torch.empty((n, k), dtype=torch.int32, device="cuda")
  nat.knn(D, k, cand)
  tab = torch.empty((n, nat.candidate_lanes(k), 2), dtype=torch.int32, device="cuda")
  nat.candidate_table(cand, D, tab)
  start = torch.from_numpy(np.stack([np.random.default_rng(50 + i).permutation(n) for i in range(P)])
                           .astype(np.int32)).cuda()
  out = {}
  for name, X in (("f64", xy), ("i32", ixy)):
    log = torch.full((P, 6001, 2), -1, dtype=torch.int64, device="cuda")
    stats = torch.zeros(3, dtype=torch.int64, device="cuda")
    cost = torch.empty(P, dtype=torch.float64, device="cuda")
    t = start.clone()
    nat.two_opt_packed_euc2d_log(t, X, tab, k, -1, log, stats, cost)
    out[name] = (t.cpu().numpy(), stats.cpu().tolist(), cost.cpu().numpy(), log.cpu().numpy()[:, :, 0])
  for a, b in zip(out["f64"], out["i32"]):
    np.testing.assert_array_equal(a, b)


def test_pass_log_records_exact_costs_after_each_move(env):
  """The logged coordinate scan: same tours as the plain launch; entry m of a tour's log is the
  cost of the tour the plain kernel leaves after m passes (one move per pass until the optimum);
  the clock never runs backwards; entries past a tour's last move stay untouched."""
  torch, nat = env
  n, P, k = 3000, 24, 8
  xy_h = np.random.default_rng(5).integers(0, 1_000_001, (n, 2)).astype(np.float64)
  xy = torch.from_numpy(xy_h).cuda()
  D = torch.empty((n, n), dtype=torch.int32, device="cuda")
  nat.dist_euc2d(xy, D)
  cand = torch.empty((n, k), dtype=torch.int32, device="cuda")
  nat.knn(D, k, cand)
  tab = torch.empty((n, nat.candidate_lanes(k), 2), dtype=torch.int32, device="cuda")
  nat.candidate_table(cand, D, tab)
  start = torch.from_numpy(np.stack([np.random.default_rng(i).permutation(n) for i in range(P)]).astype(np.int32)).cuda()
  M = 6000
  log = torch.full((P, M + 1, 2), -1, dtype=torch.int64, device="cuda")
  stats = torch.zeros(3, dtype=torch.int64, device="cuda")
  got = start.clone()
  nat.two_opt_packed_euc2d_log(got, xy, tab, k, -1, log, stats)
  plain = start.clone()
  nat.two_opt_packed_euc2d(plain, xy, tab, k, -1)
  np.testing.assert_array_equal(got.cpu().numpy(), plain.cpu().numpy())
  h = log.cpu().numpy()
  moves = (h[:, 1:, 0] >= 0).sum(axis=1)
  assert moves.sum() == int(stats[1].item()) and moves.max() < M
  cost = torch.empty(P, dtype=torch.float64, device="cuda")
  nat.tour_cost(start, D, cost)
  np.testing.assert_array_equal(h[:, 0, 0], cost.cpu().numpy().astype(np.int64))
  for m in (1, 7, 40, int(moves.min())):
    t = start.clone()
    nat.two_opt_packed_euc2d(t, xy, tab, k, m)
    nat.tour_cost(t, D, cost)
    np.testing.assert_array_equal(h[:, m, 0], cost.cpu().numpy().astype(np.int64), err_msg=f"after {m} moves")
  for p in range(P):
    c, clk = h[p, : moves[p] + 1, 0], h[p, : moves[p] + 1, 1]
    assert (np.diff(c) < 0).all() and (np.diff(clk) >= 0).all()
    assert (h[p, moves[p] + 1 :] == -1).all()
  assert nat.wall_clock_khz() > 0
