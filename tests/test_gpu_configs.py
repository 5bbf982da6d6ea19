"""The benchmark configurations at their full workloads, against the oracle.

* Config 3 (n = 1000 clustered, P = 1024, tournament 5, PMX, inversion 0.05, elite-16 2-opt
  k = 20 max_iter 10): the device loop's population and fitness after every one of 10
  generations equal the oracle GA's.
* Config 4 (256 instances, n in [50, 2000], P = 64): 4 batched generations keep every row a
  permutation, every key the row's cost and every population in survival order (its elite rows
  refined in place); the smallest, a
  middle and the largest instance equal the single-instance device loop, and the smallest the
  oracle GA.
* Config 5 (n = 10000 uniform, P = 256): scorer features, forward scores and top-8 lists on 200
  sampled cities against oracle/mlp.py, and 20 passes of learned-8 2-opt on the P tours against
  the C oracle on two rows.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_two_opt
from oracle import ga as oga
from oracle import init as oinit
from oracle import mlp
from oracle import operators as ops
from oracle.distances import compute_candidate_list, compute_distance_matrix

pytestmark = pytest.mark.gpu


def say(msg):
  print(msg, flush=True)


@pytest.fixture(scope="module")
def t():
  import tsp_gpu_clean_amd as pkg
  from tsp_gpu_clean_amd import _native

  _native.require_device()
  return pkg


def c3_coords(n=1000, seed=42):
  rng = np.random.default_rng(seed)
  centres = rng.uniform(0, 1e6, (20, 2))
  lab = rng.integers(0, 20, n)
  return np.rint(np.clip(centres[lab] + rng.normal(0, 0.02 * 1e6, (n, 2)), 0, 1e6))


def test_config3_device_loop_matches_oracle_every_generation(t):
  import torch

  from tsp_gpu_clean_amd.engine import DeviceMemeticLoop, draw_generations, upload
  from tsp_gpu_clean_amd.rng import draw_row_seeds

  n, P, G = 1000, 1024, 10
  say(f"config 3: n={n} P={P} {G} generations vs oracle")
  ctx = t.ProblemContext(t.Problem.from_coordinates("c3", c3_coords(n)), seed=42)
  d = ctx.get_cpu_distances()
  np.testing.assert_array_equal(d, compute_distance_matrix(c3_coords(n)))
  rng = np.random.default_rng(42)
  draw_row_seeds(ctx.rng, P)
  pop0 = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(rng, P))
  draws = upload(draw_generations(ctx.rng, G, P, n, 5))
  loop = DeviceMemeticLoop(ctx, P, 0.05, tournament_k=5, elite_count=16, ls_max_iterations=10, ls_candidate_k=20)
  loop.load(torch.from_numpy(pop0).cuda())
  res = oga.evolve(rng, d, pop0, max_generations=G, mutation_rate=0.05, tournament_k=5,
                   local_search=oga.LocalSearchSpec("best", 10, compute_candidate_list(d, 20)), elite_count=16,
                   patience=10**6, record=True)
  for g in range(G):
    loop.step(draws, g)
    np.testing.assert_array_equal(loop.population.cpu().numpy(), res.populations[g], err_msg=f"generation {g + 1}")
    np.testing.assert_array_equal(loop.fitness.cpu().numpy(), res.fitnesses[g])
  loop.verify()
  assert ctx.rng.bit_generator.state == rng.bit_generator.state


def test_config4_uneven_batch_invariants_and_samples(t):
  import torch

  from tsp_gpu_clean_amd.batch import BatchedMemeticGA, c4_instances
  from tsp_gpu_clean_amd.engine import DeviceMemeticLoop, draw_generations, upload
  from tsp_gpu_clean_amd.rng import draw_row_seeds

  P, G, E = 64, 4, 16
  coords = c4_instances(256, seed=7)
  ns = np.array([c.shape[0] for c in coords])
  say(f"config 4: 256 instances n in [{ns.min()}, {ns.max()}], P={P}, {G} generations")
  seeds = list(range(256))
  ga = BatchedMemeticGA(coords, population_size=P, mutation_rate=0.05, tournament_k=5, elite_count=E,
                        ls_max_iterations=10, ls_candidate_k=20, seeds=seeds)
  ga.init_random()
  d = ga.draw(G)
  best = torch.empty((G, 256), dtype=torch.int32, device="cuda")
  ga.run(d, G, best)
  torch.cuda.synchronize()
  for i in range(256):
    pop = ga.population(i).cpu().numpy()
    fit = ga.fitness(i).cpu().numpy()
    n = ns[i]
    assert (np.sort(pop, axis=1) == np.arange(n)).all(), f"instance {i}: a row is no permutation"
    np.testing.assert_array_equal(fit, ops.tour_costs(pop, ga.distances(i).cpu().numpy()), err_msg=f"instance {i}")
    # survival order, then the E elite rows refined in place (each only lower than its rank's cost)
    assert (np.diff(fit[E:]) >= 0).all() and fit[:E].max() <= fit[E], f"instance {i}: not in survival order"
  order = np.argsort(ns, kind="stable")
  for i in (int(order[0]), int(order[128]), int(order[-1])):
    say(f"config 4: instance {i} (n={ns[i]}) vs the single-instance loop")
    ctx = t.ProblemContext(t.Problem.from_coordinates(f"b{i}", coords[i]), seed=seeds[i])
    pop0 = torch.from_numpy(oinit.random_init(np.arange(ns[i]), P, draw_row_seeds(ctx.rng, P))).cuda()
    draws = upload(draw_generations(ctx.rng, G, P, int(ns[i]), 5))
    loop = DeviceMemeticLoop(ctx, P, 0.05, tournament_k=5, elite_count=E, ls_max_iterations=10, ls_candidate_k=20)
    loop.load(pop0)
    bi = torch.empty(G, dtype=torch.int32, device="cuda")
    loop.run(draws, G, bi)
    np.testing.assert_array_equal(ga.population(i).cpu().numpy(), loop.population.cpu().numpy())
    np.testing.assert_array_equal(ga.fitness(i).cpu().numpy(), loop.fitness.cpu().numpy())
    np.testing.assert_array_equal(best[:, i].cpu().numpy(), bi.cpu().numpy())
  i = int(order[0])
  say(f"config 4: instance {i} (n={ns[i]}) vs the oracle GA")
  dd = ga.distances(i).cpu().numpy()
  rng = np.random.default_rng(seeds[i])
  pop0 = oinit.random_init(np.arange(ns[i]), P, oinit.draw_row_seeds(rng, P))
  res = oga.evolve(rng, dd, pop0, max_generations=G, mutation_rate=0.05, tournament_k=5,
                   local_search=oga.LocalSearchSpec("best", 10, compute_candidate_list(dd, 20)), elite_count=E,
                   patience=10**6, record=True)
  np.testing.assert_array_equal(ga.population(i).cpu().numpy(), res.populations[-1])
  np.testing.assert_array_equal(ga.fitness(i).cpu().numpy(), res.fitnesses[-1])


def test_config5_scorer_and_learned_two_opt_at_n10000(t):
  import torch

  from tsp_gpu_clean_amd.scorer import LearnedCandidateScorer

  n, P = 10000, 256
  say(f"config 5: n={n} P={P} scorer kernels and learned-8 2-opt vs oracle")
  xy = np.random.default_rng(2024).integers(0, 1_000_001, (n, 2)).astype(np.float64)
  ctx = t.ProblemContext(t.Problem.from_coordinates("c5", xy), seed=5)
  D = compute_distance_matrix(xy)
  np.testing.assert_array_equal(ctx.get_cpu_distances(), D)
  nn32 = compute_candidate_list(D, 32)
  np.testing.assert_array_equal(ctx.get_cpu_candidate_list(32), nn32)
  sc = LearnedCandidateScorer.load()
  X = sc.features(ctx)
  z = sc.scores(X)
  sample = np.sort(np.random.default_rng(9).choice(n, 200, replace=False))
  rows = (sample[:, None] * 32 + np.arange(32)).reshape(-1)
  want_x = mlp.bf16_bits(mlp.features(xy, nn32, D))[rows]
  got_x = X.cpu().numpy().view(np.uint16)[rows]
  np.testing.assert_array_equal(got_x, want_x)
  want_z, _ = mlp.forward(want_x, sc.wbf.cpu().numpy().view(np.uint16))
  got_z = z.cpu().numpy().astype(np.float64)[rows]
  assert (np.abs(got_z - want_z) <= 1e-3 * (1 + np.abs(want_z))).all()
  cand = sc.candidates(ctx, 8).cpu().numpy()
  zz = z.cpu().numpy().reshape(n, 32)[sample].astype(np.float64)
  key = np.where(np.isnan(zz), np.inf, -zz)
  order = np.lexsort((np.broadcast_to(np.arange(32), zz.shape), key), axis=1)[:, :8]
  np.testing.assert_array_equal(cand[sample], np.take_along_axis(nn32[sample], order, axis=1))
  routes = np.stack([np.random.default_rng(100 + r).permutation(n) for r in range(P)]).astype(np.int32)
  ls = t.LocalSearchGPUTourTwoOpt(max_iterations=20, mode="best", candidate_k=8, candidate_scorer=sc)
  got = ls.improve_batch(torch.from_numpy(routes).cuda(), ctx).cpu().numpy()
  np.testing.assert_array_equal(ctx.get_gpu_learned_candidates(sc, 8)[0].cpu().numpy(), cand)
  for r in (0, P - 1):
    want, _ = c_two_opt.improve_batch(routes[r : r + 1], D, cand, "best", 20)
    np.testing.assert_array_equal(got[r : r + 1], want, err_msg=f"row {r}")
  assert (np.sort(got, axis=1) == np.arange(n)).all()
