"""GPU strategies behind the reference's strategy API, and the GA end to end, against the oracle.

Each strategy draws from ``context.rng`` exactly as the reference strategy does, so the oracle
run with a generator of the same seed must produce identical routes.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import c_two_opt
from oracle import ga as oga
from oracle import init as oinit
from oracle import operators as ops
from oracle import two_opt as o2
from oracle.distances import candidate_boundary_ties, compute_candidate_list

pytestmark = pytest.mark.gpu


def say(msg):
  print(msg, flush=True)


@pytest.fixture(scope="module")
def t():
  import tsp_gpu_clean_amd as pkg
  from tsp_gpu_clean_amd import _native

  _native.require_device()
  return pkg


def c1_problem(t, n=52, seed=42, hi=1000):
  xy = np.random.default_rng(seed).integers(0, hi + 1, (n, 2)).astype(np.float64)
  return t.Problem.from_coordinates(f"u{n}", xy)


def clustered_problem(t, n, seed=42):
  rng = np.random.default_rng(seed)
  centres = rng.uniform(0, 1e6, (20, 2))
  lab = rng.integers(0, 20, n)
  xy = np.rint(np.clip(centres[lab] + rng.normal(0, 0.02 * 1e6, (n, 2)), 0, 1e6))
  return t.Problem.from_coordinates(f"c{n}", xy)


def test_residency_shape_and_no_aliasing(t):
  import torch

  say("strategy residency")
  ctx = t.ProblemContext(c1_problem(t), seed=1)
  ls = t.LocalSearchGPUTourTwoOpt(max_iterations=5)
  host = np.stack([np.random.default_rng(i).permutation(52) for i in range(4)]).astype(np.int32)
  keep = host.copy()
  out = ls.improve_batch(host, ctx)
  assert isinstance(out, np.ndarray) and out.shape == host.shape
  np.testing.assert_array_equal(host, keep)
  dev = torch.from_numpy(host).cuda()
  out_d = ls.improve_batch(dev, ctx)
  assert isinstance(out_d, torch.Tensor) and out_d.is_cuda and out_d.data_ptr() != dev.data_ptr()
  np.testing.assert_array_equal(dev.cpu().numpy(), keep)
  np.testing.assert_array_equal(out_d.cpu().numpy(), out)
  one = ls.improve(host[0], ctx)
  assert one.shape == (52,)
  np.testing.assert_array_equal(one, out[0])
  ev = t.EvaluationGPUCost()
  assert ev.evaluate(host[0], ctx) == ops.tour_costs(host[:1], ctx.get_cpu_distances())[0]
  a = ev.assess_batch(dev, ctx, return_gpu=True)
  assert a.objectives.is_cuda and bool(a.is_feasible.all())


@pytest.mark.parametrize("k", [None, 20])
@pytest.mark.parametrize("mode", ["batch", "first", "multi_move"])
def test_local_search_strategy_matches_oracle(t, k, mode):
  say(f"LocalSearchGPUTourTwoOpt k={k} mode={mode}")
  ctx = t.ProblemContext(c1_problem(t, 300, seed=7, hi=1_000_000), seed=3)
  routes = np.stack([np.random.default_rng(i).permutation(300) for i in range(16)]).astype(np.int32)
  ls = t.LocalSearchGPUTourTwoOpt(max_iterations=10, mode=mode, candidate_k=k)
  got = ls.improve_batch(routes, ctx)
  cand = None if k is None else ctx.get_cpu_candidate_list(k)
  want = o2.improve_batch(routes, ctx.get_cpu_distances(), cand, "best" if mode == "batch" else mode, 10)
  np.testing.assert_array_equal(got, want)
  deltas, moves, passes = ls.read_stats()
  assert moves > 0 and (passes >= moves or mode == "multi_move")


def test_operator_strategies_match_oracle_with_same_generator(t):
  say("selection/crossover/mutation strategies")
  prob = c1_problem(t, 120, seed=9, hi=1_000_000)
  ctx = t.ProblemContext(prob, seed=123)
  rng = np.random.default_rng(123)
  pop = np.stack([np.random.default_rng(100 + i).permutation(120) for i in range(64)]).astype(np.int32)
  d = ctx.get_cpu_distances()
  fit = ops.tour_costs(pop, d)
  sel = t.SelectionGPUTournament(k=5).select_batch(pop, fit, 64, ctx)
  want_sel = ops.tournament_select(rng, pop, fit, 5, 64)
  np.testing.assert_array_equal(sel, want_sel)
  kids = t.CrossoverGPUPermPartiallyMapped().crossover_batch(sel, ctx)
  want_kids = ops.pmx_batch(rng, want_sel)
  np.testing.assert_array_equal(kids, want_kids)
  mut = t.MutationGPUPermInversion().mutate_batch(kids, 0.3, ctx)
  np.testing.assert_array_equal(mut, ops.inversion_batch(rng, want_kids, 0.3))
  assert ctx.rng.bit_generator.state == rng.bit_generator.state


def test_initializers_match_oracle(t):
  say("initializers")
  prob = c1_problem(t)
  ctx = t.ProblemContext(prob, seed=42)
  rng = np.random.default_rng(42)
  pop = t.InitializationGPUPermRandom(return_gpu=False).initialize(list(range(52)), 128, ctx)
  np.testing.assert_array_equal(pop, oinit.random_init(np.arange(52), 128, oinit.draw_row_seeds(rng, 128)))
  pop = t.InitializationGPUTourNearestNeighbor(k=5, return_gpu=False).initialize(list(range(52)), 128, ctx)
  cand = ctx.get_cpu_candidate_list(5)
  np.testing.assert_array_equal(pop, oinit.nn_init(np.arange(52), 128, cand, oinit.draw_row_seeds(rng, 128)))


def test_warmup_leaves_rng_and_counters_untouched(t):
  say("warmup contract")
  ctx = t.ProblemContext(c1_problem(t), seed=5)
  before = ctx.rng.bit_generator.state
  for s in (t.LocalSearchGPUTourTwoOpt(candidate_k=20), t.EvaluationGPUCost(), t.SelectionGPUTournament(),
            t.CrossoverGPUPermPartiallyMapped(), t.MutationGPUPermInversion(), t.InitializationGPUPermRandom(),
            t.InitializationGPUTourNearestNeighbor(k=5)):
    s.warmup(ctx)
    assert s.kernel_launches == 0 and s.call_count == 0 and s.h2d_bytes == 0
  assert ctx.rng.bit_generator.state == before


def _c1_ga(t, ctx, generations=50, patience=None):
  ls = t.LocalSearchGPUTourTwoOpt(mode="best", max_iterations=10, candidate_k=20)
  ev = t.EvaluationGPUCost()
  ga = t.GeneticAlgorithm(
    population_size=128, mutation_rate=0.05, seed=42,
    initialization=t.InitializationGPUTourNearestNeighbor(k=5, seed=42),
    selection=t.SelectionGPUTournament(k=5, seed=42), crossover=t.CrossoverGPUPermPartiallyMapped(seed=42),
    mutation=t.MutationGPUPermInversion(seed=42), local_search=ls, evaluation=ev, local_search_stage="none",
    pipeline=[t.PostInitLocalSearchStep(ls, ev), t.EliteOnlyLocalSearchStep(ls, ev, elite_count=16)])
  return ga, ga.evolve(ctx, list(range(ctx.dimension)), generations, patience=patience)


@pytest.mark.parametrize("n,k", [(49, 6), (400, 20), (1000, 20), (52, 5)])
def test_context_candidate_list_is_the_reference_table_on_tie_rows(t, n, k):
  """Lattice cities (k-th boundary ties on most rows): the context's device table, with its
  flagged rows rebuilt by the reference's argpartition rule, equals the oracle's table."""
  say(f"context k-NN on lattice n={n} k={k}")
  side = int(np.ceil(np.sqrt(n)))
  xy = np.array([[x, y] for x in range(side) for y in range(side)][:n], dtype=np.float64)
  ctx = t.ProblemContext(t.Problem.from_coordinates("lattice", xy), seed=1)
  d = ctx.get_cpu_distances()
  assert ctx.candidate_boundary_ties(k) == candidate_boundary_ties(d, k) > 0
  np.testing.assert_array_equal(ctx.get_cpu_candidate_list(k), compute_candidate_list(d, k))
  np.testing.assert_array_equal(ctx.get_gpu_candidate_list(k).cpu().numpy(), compute_candidate_list(d, k))


def test_local_search_strategy_beyond_the_lds_limits(t):
  """LocalSearchGPUTourTwoOpt at n = 20,000 (above the LDS-resident candidate and full scans)
  takes the workspace kernels: candidate lists k = 8 and the full neighbourhood, against the C
  oracle on one row; other modes are refused there."""
  n = 20000
  say(f"strategy beyond the LDS limits n={n}")
  xy = np.random.default_rng(20).integers(0, 1_000_001, (n, 2)).astype(np.float64)
  ctx = t.ProblemContext(t.Problem.from_coordinates("u20000", xy), seed=1)
  routes = np.stack([np.random.default_rng(i).permutation(n) for i in range(2)]).astype(np.int32)
  d = ctx.get_cpu_distances()
  for k, iters in ((8, 400), (None, 2)):
    got = t.LocalSearchGPUTourTwoOpt(max_iterations=iters, candidate_k=k).improve_batch(routes, ctx)
    want, _ = c_two_opt.improve_batch(routes[:1], d, None if k is None else compute_candidate_list(d, k), "best", iters)
    np.testing.assert_array_equal(got[:1], want)
  with pytest.raises(ValueError, match="steepest"):
    t.LocalSearchGPUTourTwoOpt(max_iterations=2, mode="first").improve_batch(routes, ctx)


def test_config1_ga_end_to_end_matches_oracle(t):
  """Config 1: 52 uniform integer cities in [0,1000]^2 (seed 42), README memetic pipeline, 50 generations."""
  say("GA config 1 end to end")
  prob = c1_problem(t)
  ctx = t.ProblemContext(prob, seed=42)
  ga, (best, stats) = _c1_ga(t, ctx)
  d = ctx.get_cpu_distances()
  # the oracle builds its own k-NN tables with the reference's argpartition rule; config 1 has
  # k-th boundary ties at k = 20, which the context rebuilds by the same rule
  c5, c20 = compute_candidate_list(d, 5), compute_candidate_list(d, 20)
  assert candidate_boundary_ties(d, 20) > 0
  np.testing.assert_array_equal(ctx.get_cpu_candidate_list(20), c20)
  np.testing.assert_array_equal(ctx.get_cpu_candidate_list(5), c5)
  rng = np.random.default_rng(42)
  pop0 = oinit.nn_init(np.arange(52), 128, c5, oinit.draw_row_seeds(rng, 128))
  res = oga.evolve(rng, d, pop0, max_generations=50, mutation_rate=0.05, tournament_k=5,
                   local_search=oga.LocalSearchSpec("best", 10, c20), post_init=True, elite_count=16)
  assert stats["convergence_history"] == res.stats["convergence_history"]
  assert stats["generations_completed"] == res.stats["generations_completed"]
  assert stats["stop_reason"] == res.stats["stop_reason"]
  np.testing.assert_array_equal(best, res.best_tour)
  assert stats["best_fitness"] == ops.tour_costs(best[None], d)[0]
  assert len(stats["convergence_history"]) == stats["generations_completed"] + 1
  assert all(a >= b for a, b in zip(stats["convergence_history"], stats["convergence_history"][1:]))


def test_device_loop_matches_step_pipeline_and_oracle(t):
  """The no-sync device loop equals the oracle GA population after every generation."""
  import torch

  from tsp_gpu_clean_amd.engine import DeviceMemeticLoop, draw_generations, upload

  say("device memetic loop vs oracle")
  n, P, G = 200, 128, 12
  prob = clustered_problem(t, n)
  ctx = t.ProblemContext(prob, seed=42)
  rng = np.random.default_rng(42)
  from tsp_gpu_clean_amd.rng import draw_row_seeds

  draw_row_seeds(ctx.rng, P)  # the device run consumes the same init draw as the oracle below
  pop0 = oinit.random_init(np.arange(n), P, oinit.draw_row_seeds(rng, P))
  draws = upload(draw_generations(ctx.rng, G, P, n, 5))
  loop = DeviceMemeticLoop(ctx, P, 0.05, tournament_k=5, elite_count=16, ls_max_iterations=10, ls_candidate_k=20)
  loop.load(torch.from_numpy(pop0).cuda())
  res = oga.evolve(rng, ctx.get_cpu_distances(), pop0, max_generations=G, mutation_rate=0.05, tournament_k=5,
                   local_search=oga.LocalSearchSpec("best", 10, ctx.get_cpu_candidate_list(20)), elite_count=16,
                   patience=10**6, record=True)
  for g in range(G):
    loop.step(draws, g)
    np.testing.assert_array_equal(loop.population.cpu().numpy(), res.populations[g], err_msg=f"generation {g + 1}")
    np.testing.assert_array_equal(loop.fitness.cpu().numpy(), res.fitnesses[g])
  loop.verify()
  assert not loop.sync.view(-1, 528)[:, list(range(0, 256, 16)) + list(range(1, 256, 16)) + [512, 513, 514]].any().item()
  assert ctx.rng.bit_generator.state == rng.bit_generator.state


def test_device_loop_graph_replays_match_eager(t):
  """A captured run of generations replayed several times (eager work and syncs in between)
  gives the eager result every time."""
  import torch

  from tsp_gpu_clean_amd.engine import DeviceMemeticLoop, draw_generations, upload
  from tsp_gpu_clean_amd.rng import draw_row_seeds

  say("device loop hipGraph replays")
  n, P, G = 150, 64, 8
  ctx = t.ProblemContext(clustered_problem(t, n, seed=3), seed=11)
  seeds = draw_row_seeds(ctx.rng, P)
  pop0 = torch.from_numpy(oinit.random_init(np.arange(n), P, seeds)).cuda()
  draws = upload(draw_generations(ctx.rng, G, P, n, 5))
  loop = DeviceMemeticLoop(ctx, P, 0.2, tournament_k=5, elite_count=8, ls_max_iterations=5, ls_candidate_k=10)
  best = torch.empty(G, dtype=torch.int32, device="cuda")
  loop.load(pop0)
  loop.run(draws, G, best)
  want_pop = loop.population.cpu().numpy()
  want_fit = loop.fitness.cpu().numpy()
  want_best = best.cpu().numpy()
  loop.load(pop0)
  graph = loop.capture(draws, G, best)
  for rep in range(3):
    loop.load(pop0)
    best.zero_()
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(loop.population.cpu().numpy(), want_pop, err_msg=f"replay {rep}")
    np.testing.assert_array_equal(loop.fitness.cpu().numpy(), want_fit)
    np.testing.assert_array_equal(best.cpu().numpy(), want_best)
    _ = (loop.population.float() * 2).sum().item()  # unrelated eager work between replays
  loop.verify()


@pytest.mark.parametrize("n", [5000, 8000])
def test_local_search_strategy_large_n_computes_distances_from_coordinates(t, n):
  """Integer instances: the strategy's candidate scan reads coordinates (node records in LDS at
  n = 5000; above, the instance renumbered in Hilbert order inside the call), with the matrix
  kernel's tours."""
  import torch

  from tsp_gpu_clean_amd import _native

  say(f"LocalSearchGPUTourTwoOpt n={n} k=8 (coordinate candidate scan)")
  rng = np.random.default_rng(31)
  xy = rng.integers(0, 1_000_001, (n, 2)).astype(np.float64)
  ctx = t.ProblemContext(t.Problem.from_coordinates(f"u{n}", xy), seed=2)
  routes = np.stack([rng.permutation(n) for _ in range(4)]).astype(np.int32)
  ls = t.LocalSearchGPUTourTwoOpt(max_iterations=40, mode="best", candidate_k=8)
  got = ls.improve_batch(routes, ctx)
  want = torch.from_numpy(routes).cuda()
  _native.two_opt_packed(want, ctx.get_gpu_distances(), ctx.get_gpu_candidate_table(8), 8, 40, None)
  np.testing.assert_array_equal(got, want.cpu().numpy())
  assert ls.read_stats()[1] > 0


def test_local_search_strategy_many_tours_coordinate_scan_matches_c_oracle(t):
  """n = 1000 with more tours than CUs: the strategy takes the coordinate candidate scan (512-thread
  workgroups); tours and counters equal the C oracle's."""
  import torch

  from tsp_gpu_clean_amd.context import coordinate_candidates_pay

  say("LocalSearchGPUTourTwoOpt n=1000 P=300 k=20 (many-tour coordinate scan)")
  P = max(300, torch.cuda.get_device_properties(0).multi_processor_count + 1)
  assert coordinate_candidates_pay(1000, P)
  rng = np.random.default_rng(41)
  xy = rng.integers(0, 1_000_001, (1000, 2)).astype(np.float64)
  ctx = t.ProblemContext(t.Problem.from_coordinates("u1000", xy), seed=2)
  routes = np.stack([rng.permutation(1000) for _ in range(P)]).astype(np.int32)
  ls = t.LocalSearchGPUTourTwoOpt(max_iterations=15, mode="best", candidate_k=20)
  got = ls.improve_batch(routes, ctx)
  rows = [0, P // 2, P - 1]
  cand = ctx.get_gpu_candidate_list(20).cpu().numpy()  # the table the strategy scans
  want, _ = c_two_opt.improve_batch(routes[rows], ctx.get_cpu_distances(), cand, "best", 15)
  np.testing.assert_array_equal(got[rows], want)


def test_chapter4_operators_match_oracle_with_same_generator(t):
  """Order crossover and swap mutation strategies: the same draws and children as the oracle."""
  say("OX / swap strategies")
  prob = c1_problem(t, 120, seed=9, hi=1_000_000)
  ctx = t.ProblemContext(prob, seed=321)
  rng = np.random.default_rng(321)
  pop = np.stack([np.random.default_rng(200 + i).permutation(120) for i in range(65)]).astype(np.int32)
  kids = t.CrossoverGPUPermOrder().crossover_batch(pop, ctx)
  want = ops.ox_batch(rng, pop)
  np.testing.assert_array_equal(kids, want)
  mut = t.MutationGPUPermSwap().mutate_batch(kids, 0.4, ctx)
  np.testing.assert_array_equal(mut, ops.swap_batch(rng, want, 0.4))
  child = t.CrossoverGPUPermOrder().crossover(pop[0], pop[1], ctx)
  cuts = ops.sample_distinct_pairs(rng, 1, 0, 120)[0]
  np.testing.assert_array_equal(child, ops.ox_pair(pop[0], pop[1], min(cuts), max(cuts)))
  assert ctx.rng.bit_generator.state == rng.bit_generator.state


def test_chapter4_ga_end_to_end_matches_oracle(t):
  """The config-1 memetic pipeline with OX and swap in place of PMX and inversion."""
  say("GA with OX + swap end to end")
  prob = c1_problem(t)
  ctx = t.ProblemContext(prob, seed=42)
  ls = t.LocalSearchGPUTourTwoOpt(mode="best", max_iterations=10, candidate_k=20)
  ev = t.EvaluationGPUCost()
  ga = t.GeneticAlgorithm(
    population_size=128, mutation_rate=0.1, seed=42,
    initialization=t.InitializationGPUTourNearestNeighbor(k=5, seed=42),
    selection=t.SelectionGPUTournament(k=5, seed=42), crossover=t.CrossoverGPUPermOrder(seed=42),
    mutation=t.MutationGPUPermSwap(seed=42), local_search=ls, evaluation=ev, local_search_stage="none",
    pipeline=[t.PostInitLocalSearchStep(ls, ev), t.EliteOnlyLocalSearchStep(ls, ev, elite_count=16)])
  best, stats = ga.evolve(ctx, list(range(52)), 30)
  rng = np.random.default_rng(42)
  pop0 = oinit.nn_init(np.arange(52), 128, ctx.get_cpu_candidate_list(5), oinit.draw_row_seeds(rng, 128))
  res = oga.evolve(rng, ctx.get_cpu_distances(), pop0, max_generations=30, mutation_rate=0.1, tournament_k=5,
                   local_search=oga.LocalSearchSpec("best", 10, ctx.get_cpu_candidate_list(20)), post_init=True,
                   elite_count=16, crossover="ox", mutation="swap")
  assert stats["convergence_history"] == res.stats["convergence_history"]
  assert stats["stop_reason"] == res.stats["stop_reason"]
  np.testing.assert_array_equal(best, res.best_tour)


@pytest.mark.parametrize("name", ["burma14", "gr17", "bayg29", "att48", "berlin52"])
def test_reference_fixture_instances(t, name):
  """The reference's fixture instances (GEO, EXPLICIT lower/upper rows, ATT, EUC_2D) run through the
  memetic pipeline: device distances equal the oracle's, and the GA reaches the published optimum
  (problems.json) on the small ones and never goes below it."""
  import json
  import os

  from oracle.distances import compute_distance_matrix
  from tests.conftest import GOLDEN

  say(f"fixture {name}")
  with open(os.path.join(GOLDEN, "problems.json"), encoding="utf-8") as fh:
    optimum = {p["name"]: p["optimal"] for p in json.load(fh)["problems"]}[name]
  prob = t.load_tsplib_file(os.path.join(GOLDEN, f"{name}.tsp"))
  ctx = t.ProblemContext(prob, seed=7)
  want = prob.distances if prob.edge_type == "EXPLICIT" else compute_distance_matrix(prob.coordinates, prob.edge_type)
  np.testing.assert_array_equal(ctx.get_cpu_distances(), want)
  ls = t.LocalSearchGPUTourTwoOpt(mode="best", max_iterations=50)
  ev = t.EvaluationGPUCost()
  ga = t.GeneticAlgorithm(
    population_size=128, mutation_rate=0.2, seed=7,
    initialization=t.InitializationGPUPermRandom(seed=7), selection=t.SelectionGPUTournament(k=5, seed=7),
    crossover=t.CrossoverGPUPermOrder(seed=7), mutation=t.MutationGPUPermInversion(seed=7), local_search=ls,
    evaluation=ev, local_search_stage="none",
    pipeline=[t.PostInitLocalSearchStep(ls, ev), t.EliteOnlyLocalSearchStep(ls, ev, elite_count=32)])
  best, stats = ga.evolve(ctx, list(range(prob.dimension)), 150, patience=60)
  cost = ops.tour_costs(best[None], want)[0]
  assert cost == stats["best_fitness"] and cost >= optimum
  if prob.dimension <= 29:
    assert cost == optimum, (name, cost, optimum)
  else:
    assert cost <= 1.03 * optimum, (name, cost, optimum)


def test_runner_record_for_reference_config(t):
  """The reference's FullGPUModular configuration (tests/golden/algorithms.json) on berlin52 for
  three seeded repetitions: a complete summary record whose runs replay exactly."""
  import json
  import os

  from tests.conftest import GOLDEN
  from tests.test_runner_cpu import REFERENCE_KEYS
  from tsp_gpu_clean_amd.runner import build_algorithm, run_repetitions

  say("runner record FullGPUModular berlin52")
  with open(os.path.join(GOLDEN, "algorithms.json"), encoding="utf-8") as fh:
    cfg = json.load(fh)
  prob = t.load_tsplib_file(os.path.join(GOLDEN, "berlin52.tsp"))
  spec = cfg["algorithm_configs"]["FullGPUModular"]
  factory = lambda seed: build_algorithm(spec, cfg["default_params"], seed=seed)  # noqa: E731
  rec = run_repetitions("FullGPUModular", factory, prob, max_generations=40, optimal_cost=7542, repetitions=3)
  assert set(rec) == REFERENCE_KEYS
  assert rec["successful_runs"] == 3 and rec["raw_seeds"] == [42, 43, 44]
  assert all(c >= 7542 for c in rec["raw_costs"]) and rec["best_cost"] == min(rec["raw_costs"])
  assert rec["mean_kernels"] > 0 and rec["mean_phase_timings"] and rec["mean_time"] > 0
  tour, stats = factory(43).evolve(t.ProblemContext(prob, seed=43), list(range(52)), 40, optimal_cost=7542,
                                   patience=int(2 * 52 ** 0.5))
  assert stats["best_fitness"] == rec["raw_costs"][1]


def test_reference_cpu_hybrid_and_gpu_configs_agree(t):
  """The reference's four algorithm configurations (CPU, HybridNaive, HybridOptimized,
  FullGPUModular in tests/golden/algorithms.json) differ only in class names and residency; with
  one seed they run the same GA, so their histories and best tours are identical."""
  import json
  import os

  from tests.conftest import GOLDEN
  from tsp_gpu_clean_amd.runner import build_algorithm

  say("reference CPU / Hybrid* / FullGPUModular configs on berlin52")
  with open(os.path.join(GOLDEN, "algorithms.json"), encoding="utf-8") as fh:
    cfg = json.load(fh)
  prob = t.load_tsplib_file(os.path.join(GOLDEN, "berlin52.tsp"))
  runs = {}
  for name in ("CPU", "HybridNaive", "HybridOptimized", "FullGPUModular"):
    ga = build_algorithm(cfg["algorithm_configs"][name], cfg["default_params"], seed=42)
    tour, stats = ga.evolve(t.ProblemContext(prob, seed=42), list(range(52)), 25, optimal_cost=7542, patience=10**6)
    runs[name] = (np.asarray(tour.cpu() if hasattr(tour, "cpu") else tour), stats["convergence_history"])
  want_tour, want_hist = runs["CPU"]
  for name, (tour, hist) in runs.items():
    assert hist == want_hist, name
    np.testing.assert_array_equal(tour, want_tour, err_msg=name)


def _memetic_ga(t, P, elite, *, device_loop, chunk=8, cpu_named=False, post_init=True, ls_k=20, rate=0.05):
  if cpu_named:
    ls = t.LocalSearchCPUTourTwoOpt(mode="best", max_iterations=10, candidate_k=ls_k)
    ev, sel = t.EvaluationCPUCost(), t.SelectionCPUTournament(k=5)
    init, cx, mut = t.InitializationCPUTourNearestNeighbor(k=5), t.CrossoverCPUPermPartiallyMapped(), \
      t.MutationCPUPermInversion()
  else:
    ls = t.LocalSearchGPUTourTwoOpt(mode="best", max_iterations=10, candidate_k=ls_k)
    ev, sel = t.EvaluationGPUCost(return_gpu=True), t.SelectionGPUTournament(k=5)
    init, cx, mut = t.InitializationGPUTourNearestNeighbor(k=5), t.CrossoverGPUPermPartiallyMapped(), \
      t.MutationGPUPermInversion()
  steps = ([t.PostInitLocalSearchStep(ls, ev)] if post_init else []) + \
    ([t.EliteOnlyLocalSearchStep(ls, ev, elite_count=elite)] if elite else [])
  extra = {} if chunk is None else {"device_loop_chunk": chunk}
  return t.GeneticAlgorithm(population_size=P, mutation_rate=rate, seed=42, initialization=init, selection=sel,
                            crossover=cx, mutation=mut, evaluation=ev, local_search=ls, local_search_stage="none",
                            pipeline=steps, device_loop=device_loop, **extra)


def _host(x):
  return np.asarray(x.cpu().numpy() if hasattr(x, "cpu") else x)


def _same_state(a, b) -> bool:
  """Generator states compared field by field (some bit generators hold arrays)."""
  if isinstance(a, dict):
    return isinstance(b, dict) and a.keys() == b.keys() and all(_same_state(a[k], b[k]) for k in a)
  if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
    return bool(np.array_equal(a, b))
  return a == b


@pytest.mark.parametrize("case", ["patience_stop", "optimal_stop", "remainder_no_stop", "no_elite", "cpu_named",
                                  "coordinates", "default_chunk", "philox_rng", "cluster_timeout"])
def test_evolve_device_loop_equals_step_pipeline(t, case, monkeypatch):
  """evolve(device_loop=True) runs graph-captured chunks of generations and must end exactly where
  the step pipeline ends: history, stop, best tour, final population, fitness and generator.  The
  chunk is chosen so that a stop falls inside a chunk (the rewind path) and a run without a stop
  ends in a partial chunk (the eager remainder).  In "cluster_timeout" the host raises the error
  word of an elite cluster before some eager elite searches (the cluster then leaves at once, as
  after a barrier timeout): the chunks concerned are run again and the result is unchanged."""
  import torch

  from tsp_gpu_clean_amd import _native
  from tsp_gpu_clean_amd.engine import device_route_plan

  forced = {"calls": 0, "raised": 0}
  if case == "cluster_timeout":
    real = _native.two_opt_elite

    def raising(tours, D, table, k, iters, stats, cost, sync, best=None, ixy=None):
      # every 4th elite search, eager or captured (a captured one raises the word in every replay)
      forced["calls"] += 1
      if forced["calls"] % 4 == 3:
        e = _native.ELITE_ERROR_WORD
        sync.view(-1, _native.ELITE_SYNC_WORDS)[0, e : e + 1].fill_(1)  # a fill kernel: capturable
        forced["raised"] += 1
        forced["captured"] = forced.get("captured", 0) + int(torch.cuda.is_current_stream_capturing())
      return real(tours, D, table, k, iters, stats, cost, sync, best, ixy=ixy)

  n, P, gens, elite, kw, opt, patience = 120, 64, 40, 8, {}, None, 3
  if case == "remainder_no_stop":
    gens, patience = 27, 10**6
  elif case == "no_elite":
    elite, kw, gens, patience = None, {"post_init": False}, 21, 10**6
  elif case == "cpu_named":
    kw = {"cpu_named": True}
  elif case == "coordinates":
    n, P, gens, patience = 4500, 16, 6, 10**6
  elif case == "default_chunk":  # the constructor's chunk: eager first chunk, captured replays, partial last
    n, P, gens, patience = 150, 64, 90, 10**6
  elif case == "philox_rng":  # a non-PCG64 context generator takes the per-call NumPy draws
    gens = 30
  elif case == "cluster_timeout":  # eager first chunk, captured replays, eager partial last chunk
    gens, patience = 27, 10**6
  prob = clustered_problem(t, n, seed=7)
  if case == "optimal_stop":
    ga = _memetic_ga(t, P, elite, device_loop=False)
    _, s = ga.evolve(t.ProblemContext(prob, seed=3), list(range(n)), gens, patience=10**6)
    opt = s["convergence_history"][len(s["convergence_history"]) // 2] / 1.005  # stops within 1% mid-run
    patience = 10**6
  say(f"evolve device loop vs step pipeline: {case} n={n} P={P} gens={gens}")
  def context():
    c = t.ProblemContext(prob, seed=3)
    if case == "philox_rng":
      c.rng = np.random.Generator(np.random.Philox(3))
    return c

  ga0 = _memetic_ga(t, P, elite, device_loop=False, **kw)
  c0 = context()
  b0, s0 = ga0.evolve(c0, list(range(n)), gens, optimal_cost=opt, patience=patience)
  done = s0["generations_completed"]
  if case == "default_chunk":
    chunk = None
  elif s0["stopped_early"]:
    chunk = next(c for c in (6, 8, 10, 14) if done % c)
  else:
    chunk = next(c for c in (4, 8, 10) if done % c)
  say(f"  step pipeline: {done} generations, {s0['stop_reason']}; device chunk {chunk}")
  ga1 = _memetic_ga(t, P, elite, device_loop=True, chunk=chunk, **kw)
  c1 = context()
  assert device_route_plan(ga1, c1, np.zeros((P, n), np.int32)) is not None
  if case == "cluster_timeout":
    monkeypatch.setattr(_native, "two_opt_elite", raising)
  b1, s1 = ga1.evolve(c1, list(range(n)), gens, optimal_cost=opt, patience=patience)
  if case == "cluster_timeout":
    assert forced.get("captured", 0) >= 1 and forced["raised"] >= 3, forced
    assert ga1.device_loop_recoveries >= 3, (forced, ga1.device_loop_recoveries)
  assert "device_loop" in s1["phase_timings"] and "device_loop" not in s0["phase_timings"]
  for key in ("convergence_history", "generations_completed", "stop_reason", "stopped_early", "best_fitness",
              "initial_fitness"):
    assert s1[key] == s0[key], key
  if case in ("optimal_stop",):
    assert s0["stop_reason"] == "within_1pct_optimal"
  np.testing.assert_array_equal(_host(b1), _host(b0))
  assert type(ga1.final_population) is type(ga0.final_population)
  assert type(ga1.final_fitness) is type(ga0.final_fitness)
  np.testing.assert_array_equal(_host(ga1.final_population), _host(ga0.final_population))
  np.testing.assert_array_equal(_host(ga1.final_fitness), _host(ga0.final_fitness))
  assert _same_state(c1.rng.bit_generator.state, c0.rng.bit_generator.state)
