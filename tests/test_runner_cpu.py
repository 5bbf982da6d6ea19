"""The summary record (restating src/benchmarking/runner/algorithm_runner.py:75-442 for TSP) on
hand-made run results, and config parsing without a device."""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

REFERENCE_KEYS = {
  "algorithm", "repetitions", "successful_runs", "mean_time", "std_time", "min_time", "max_time", "mean_cost",
  "std_cost", "best_cost", "worst_cost", "mean_initial_cost", "std_initial_cost", "mean_gap", "std_gap", "best_gap",
  "mean_improvement", "mean_generations", "std_generations", "min_generations", "max_generations", "mean_h2d_mb",
  "mean_d2h_mb", "total_transfer_mb", "mean_kernels", "best_tour", "raw_best_tours", "problem_type", "edge_type",
  "raw_times", "raw_initial_costs", "raw_costs", "raw_gaps", "raw_generations", "raw_stop_reasons", "raw_seeds",
  "raw_phase_timings", "mean_phase_timings", "phase_share_pct", "phase_share_cv_pct",
}


def _results():
  return {
    "times": [1.0, 2.0, float("nan"), 3.0], "initial_costs": [200.0, 100.0, float("nan"), 400.0],
    "final_costs": [110.0, 100.0, float("nan"), 120.0], "gaps": [10.0, 0.0, float("nan"), 20.0],
    "generations": [5, 7, 0, 9], "stop_reasons": ["a", "b", "failed", "c"], "seeds": [42, 43, 44, 45],
    "h2d_bytes": [1e6, 3e6, 0, 2e6], "d2h_bytes": [0, 0, 0, 4e6], "kernel_launches": [10, 20, 0, 30],
    "phase_timings": [{"ls": {"total_ms": 500.0, "calls": 5}}, {"ls": {"total_ms": 1000.0, "calls": 5}}, {},
                      {"ls": {"total_ms": 600.0, "calls": 2}}],
    "best_tours": [[0, 1, 2], [0, 2, 1], None, [1, 0, 2]],
  }


def test_summary_record_values():
  from tsp_gpu_clean_amd import Problem
  from tsp_gpu_clean_amd.runner import compute_series_stats, summary_payload

  prob = Problem.from_coordinates("p", np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]))
  rec = summary_payload(_results(), prob, "algo", has_known_optimal=True)
  assert set(rec) == REFERENCE_KEYS
  assert rec["repetitions"] == 4 and rec["successful_runs"] == 3
  assert rec["mean_time"] == pytest.approx(2.0) and rec["std_time"] == pytest.approx(1.0)  # ddof = 1
  assert (rec["best_cost"], rec["worst_cost"]) == (100.0, 120.0)
  assert rec["best_tour"] == [0, 2, 1]  # the run with the lowest final cost
  assert rec["mean_gap"] == pytest.approx(10.0) and rec["best_gap"] == 0.0
  assert rec["mean_improvement"] == pytest.approx((45.0 + 0.0 + 70.0) / 3)
  assert rec["mean_h2d_mb"] == pytest.approx(1.5) and rec["mean_d2h_mb"] == pytest.approx(1.0)  # over 4 runs
  assert rec["mean_kernels"] == pytest.approx(15.0)
  assert rec["raw_seeds"] == [42, 43, 45] and rec["raw_stop_reasons"] == ["a", "b", "c"]
  assert rec["mean_phase_timings"]["ls"]["total_ms"] == pytest.approx(700.0)
  assert rec["phase_share_pct"]["ls"] == pytest.approx((50.0 + 50.0 + 20.0) / 3)
  assert compute_series_stats(np.array([np.nan, np.nan]), 2) == (0.0, 0.0, 0.0, 0.0)
  assert compute_series_stats(np.array([3.0]), 1) == (3.0, 0.0, 3.0, 3.0)
  rec2 = summary_payload(_results(), prob, "algo", has_known_optimal=False)
  assert rec2["mean_gap"] is None and rec2["raw_gaps"] is None


def test_record_json_and_configs(tmp_path):
  from tsp_gpu_clean_amd import Problem
  from tsp_gpu_clean_amd.runner import build_algorithm, save_record, summary_payload

  prob = Problem.from_coordinates("p", np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]))
  path = tmp_path / "rec.json"
  save_record(str(path), summary_payload(_results(), prob, "algo"))
  loaded = json.loads(path.read_text())
  assert set(loaded) == REFERENCE_KEYS and loaded["raw_gaps"] == [10.0, 0.0, 20.0]
  with open(os.path.join(GOLDEN, "algorithms.json"), encoding="utf-8") as fh:
    cfg = json.load(fh)
  ga = build_algorithm(cfg["algorithm_configs"]["FullGPUModular"], cfg["default_params"])
  assert type(ga.crossover).__name__ == "CrossoverGPUPermOrder" and ga.population_size == 32
  assert ga.selection.k == 3 and ga.local_search.max_iterations == 5 and ga.local_search.candidate_k == 20
  # the reference's CPU-named classes: the same device kernels, the CPU constructors' defaults
  cpu = build_algorithm(cfg["algorithm_configs"]["CPU"], cfg["default_params"])
  assert type(cpu.local_search).__name__ == "LocalSearchCPUTourTwoOpt" and cpu.local_search.mode == "best"
  assert type(cpu.selection).__name__ == "SelectionCPUTournament" and cpu.selection.k == 3
  assert type(cpu.evaluation).__name__ == "EvaluationCPUCost" and not cpu.evaluation.return_gpu
  for name in ("HybridNaive", "HybridOptimized"):
    build_algorithm(cfg["algorithm_configs"][name], cfg["default_params"])
  with pytest.raises(KeyError):  # classes outside the TSP hot path stay unregistered
    build_algorithm({"strategies": {"selection": {"class": "SelectionCPURouletteWheel"}}}, cfg["default_params"])


def test_cpu_named_classes_follow_the_cpu_signatures():
  import tsp_gpu_clean_amd as t

  assert t.SelectionCPUTournament().k == 3 and t.InitializationCPUTourNearestNeighbor().k == 20
  assert not t.InitializationCPUPermRandom().return_gpu
  with pytest.raises(ValueError):
    t.LocalSearchCPUTourTwoOpt(mode="batch")  # GPU-only mode name
  with pytest.raises(ValueError):
    t.LocalSearchCPUTourTwoOpt(max_iterations=-1)
  with pytest.raises(NotImplementedError):
    t.EvaluationCPUCost(decoder=object())
  for name in ("SelectionCPUTournament", "CrossoverCPUPermOrder", "CrossoverCPUPermPartiallyMapped",
               "MutationCPUPermSwap", "MutationCPUPermInversion", "InitializationCPUPermRandom",
               "InitializationCPUTourNearestNeighbor", "EvaluationCPUCost", "LocalSearchCPUTourTwoOpt"):
    assert t.STRATEGY_REGISTRY[name] is getattr(t, name)


def _route_ga(t, P=64, stage="none", cx=None, mut=None, ls_kw=None, extra=None, elite=True):
  ls = t.LocalSearchGPUTourTwoOpt(**({"mode": "best", "candidate_k": 20} | (ls_kw or {})))
  ev = t.EvaluationGPUCost(return_gpu=True)
  steps = [t.PostInitLocalSearchStep(ls, ev)] + ([t.EliteOnlyLocalSearchStep(ls, ev, 8)] if elite else [])
  return t.GeneticAlgorithm(population_size=P, initialization=t.InitializationGPUPermRandom(),
                            selection=t.SelectionGPUTournament(), crossover=cx or t.CrossoverGPUPermPartiallyMapped(),
                            mutation=mut or t.MutationGPUPermInversion(), evaluation=ev, local_search=ls,
                            local_search_stage=stage, pipeline=steps + (extra or []), device_loop=True)


def test_device_route_plan_accepts_only_the_device_loop_pipeline():
  """The graph-captured route is taken only for the exact generation the device loop runs; every
  other pipeline, strategy or shape keeps the step pipeline (plan None)."""
  import tsp_gpu_clean_amd as t
  from tsp_gpu_clean_amd.engine import GENERATION_MAX_N, device_route_plan

  xy = np.random.default_rng(1).integers(0, 1000, (100, 2)).astype(np.float64)
  ctx = t.ProblemContext(t.Problem.from_coordinates("u100", xy), seed=1)
  pop = np.zeros((64, 100), np.int32)
  plan = device_route_plan(_route_ga(t), ctx, pop)
  assert plan is not None and (plan["P"], plan["k"], plan["elite"], plan["iters"], plan["cand_k"]) == (64, 5, 8, 10, 20)
  assert device_route_plan(_route_ga(t, elite=False), ctx, pop)["elite"] is None
  rejected = {
    "stage": _route_ga(t, stage="post_mutation"),
    "order crossover": _route_ga(t, cx=t.CrossoverGPUPermOrder()),
    "swap mutation": _route_ga(t, mut=t.MutationGPUPermSwap()),
    "first improvement": _route_ga(t, ls_kw={"mode": "first"}),
    "full neighbourhood": _route_ga(t, ls_kw={"candidate_k": None}),
    "candidate k > 32": _route_ga(t, ls_kw={"candidate_k": 40}),
    "no passes": _route_ga(t, ls_kw={"max_iterations": 0}),
    "periodic step": _route_ga(t, extra=[t.PeriodicLocalSearchStep(t.LocalSearchGPUTourTwoOpt(), 5, "offspring")]),
  }
  for why, ga in rejected.items():
    assert device_route_plan(ga, ctx, pop) is None, why
  assert device_route_plan(_route_ga(t, P=63), ctx, np.zeros((63, 100), np.int32)) is None  # odd population
  assert device_route_plan(_route_ga(t), ctx, np.zeros((64, 99), np.int32)) is None  # not whole-instance tours
  no_rng = t.ProblemContext(t.Problem.from_coordinates("u100", xy), seed=1)
  no_rng.rng = None
  assert device_route_plan(_route_ga(t), no_rng, pop) is None
  big = np.random.default_rng(2).integers(0, 10**6, (GENERATION_MAX_N + 1, 2)).astype(np.float64)
  bctx = t.ProblemContext(t.Problem.from_coordinates("big", big), seed=1)
  assert device_route_plan(_route_ga(t), bctx, np.zeros((64, GENERATION_MAX_N + 1), np.int32)) is None
  with pytest.raises(ValueError):
    t.GeneticAlgorithm(initialization=None, selection=None, crossover=None, mutation=None, evaluation=None,
                       device_loop=True, device_loop_chunk=7)


@pytest.mark.parametrize("P,n,k,G,pre", [(1024, 1000, 5, 3, 0), (64, 120, 5, 4, 1), (7, 5, 3, 5, 3), (2, 2, 1, 3, 0),
                                         (1, 4, 2, 2, 1), (33, 2000, 20, 2, 2), (16, 4500, 5, 1, 5)])
def test_native_generation_draws_equal_numpy_calls(P, n, k, G, pre):
  """tspmi_draw_generations (host C++) gives the NumPy Generator(PCG64) calls' values, end state
  and per-generation states, from fresh states and from ones holding a buffered 32-bit half."""
  from tsp_gpu_clean_amd.engine import draw_generations, draw_generations_numpy

  for seed in (0, 42, 2**63 + 5):
    a, b = np.random.default_rng(seed), np.random.default_rng(seed)
    for r in (a, b):
      r.integers(0, 1000, size=pre, dtype=np.int32)  # odd counts leave has_uint32 = 1
      r.random(pre)
    sa, sb = [], []
    da = draw_generations(a, G, P, n, k, sa)
    db = draw_generations_numpy(b, G, P, n, k, sb)
    for name in ("tour_cand", "cut_lo", "cut_hi", "probs", "pos1", "pos2"):
      np.testing.assert_array_equal(getattr(da, name), getattr(db, name), err_msg=name)
    assert a.bit_generator.state == b.bit_generator.state
    assert sa == sb
    assert a.integers(0, 2**31 - 1, 8).tolist() == b.integers(0, 2**31 - 1, 8).tolist()


def test_native_generation_draws_reject_bad_shapes():
  """tspmi_draw_generations refuses shapes the reference calls could not make (no population,
  fewer than two cities, no tournament candidates) and leaves the generator untouched."""
  from tsp_gpu_clean_amd import _native

  rng = np.random.default_rng(9)
  before = rng.bit_generator.state
  for G, P, n, k in ((1, 0, 10, 5), (1, 8, 1, 5), (1, 8, 10, 0), (-1, 8, 10, 5)):
    with pytest.raises(ValueError):
      _native.draw_generations_pcg64(before, G, P, n, k)
  assert rng.bit_generator.state == before
  new, cand, *_ = _native.draw_generations_pcg64(before, 0, 8, 10, 5)
  assert new == before and cand.shape == (0, 8, 5)


def test_reference_knn_rows_equals_the_reference_rule_on_tie_rows():
  """The host rebuild of flagged k-NN rows (context.reference_knn_rows) equals the oracle's
  full-table restatement of matrix.py:140-211 row for row, on tie-heavy lattices."""
  from oracle.distances import candidate_boundary_ties, compute_candidate_list
  from tests.conftest import lattice_matrix
  from tsp_gpu_clean_amd.context import reference_knn_rows

  for n, k in ((49, 6), (400, 20), (1000, 20), (52, 5)):
    d = lattice_matrix(n)
    assert candidate_boundary_ties(d, k) > 0
    rows = np.sort(np.random.default_rng(n).choice(n, size=min(n, 37), replace=False))
    np.testing.assert_array_equal(reference_knn_rows(d[rows], rows, min(k, n - 1)), compute_candidate_list(d, k)[rows])


def test_evaluation_gpu_cost_takes_the_reference_keywords():
  """cost_evaluation.py:168-176: a config entry with the VRP keywords at their TSP defaults builds;
  a VRP decoder or penalty policy raises NotImplementedError; a bad split mode is a ValueError."""
  import tsp_gpu_clean_amd as t
  from tsp_gpu_clean_amd.runner import _strategy

  ev = _strategy({"class": "EvaluationGPUCost",
                  "params": {"return_gpu": True, "threads_per_block": 200, "decoder": None, "split_mode": "greedy",
                             "penalty_policy": None}}, 3)
  assert isinstance(ev, t.EvaluationGPUCost) and ev.return_gpu and ev.split_mode == "greedy"
  assert ev.threads_per_block == 128
  assert t.EvaluationGPUCost(False, 256, None, None, "prins", None).threads_per_block == 256
  with pytest.raises(NotImplementedError):
    t.EvaluationGPUCost(decoder=object())
  with pytest.raises(NotImplementedError):
    t.EvaluationGPUCost(penalty_policy=object())
  with pytest.raises(ValueError):
    t.EvaluationGPUCost(split_mode="bogus")
  with pytest.raises(ValueError):
    t.EvaluationCPUCost(split_mode="bogus")


def test_explicit_distances_are_checked_before_the_int32_cast():
  """Values that an int32 cast would wrap or truncate are rejected before the cast."""
  from tsp_gpu_clean_amd.context import MAX_DISTANCE, explicit_int32_distances

  ok = np.array([[0, 5], [5, 0]], dtype=np.int64)
  assert explicit_int32_distances(ok).dtype == np.int32
  assert explicit_int32_distances(ok.astype(np.float64)).tolist() == [[0, 5], [5, 0]]
  for bad in (np.array([[0, 2**31], [2**31, 0]], dtype=np.int64),       # would wrap negative
              np.array([[0, 2**32 + 3], [2**32 + 3, 0]], dtype=np.int64),  # would wrap to 3
              np.array([[0, MAX_DISTANCE + 1], [MAX_DISTANCE + 1, 0]]),
              np.array([[0.0, 1.5], [1.5, 0.0]]),                          # would truncate to 1
              np.array([[0.0, np.nan], [np.nan, 0.0]]),
              np.array([[0, -1], [-1, 0]])):
    with pytest.raises(ValueError):
      explicit_int32_distances(bad)


def test_bench_diag_lane_counts_match_brute_force():
  """The bench's per-pass lane counts of the diagonal scan against a direct count over (wave, slot,
  row): slot q of wave w is live at row i >= 1 while i < (n-2) - off(w, q); the mixed scan moves 13
  rows in 16 of each phase to coordinates without changing the total."""
  import bench

  def off(w, q):
    return 64 * (4 * q + ((3 - w) if q & 1 else w))

  for n in (4, 63, 66, 257, 700, 1000, 1026):
    live = sum(1 for w in range(4) for q in range(4) for i in range(n - 2)
               if i == 0 or i < (n - 2) - off(w, q))
    g, c = bench.diag_lanes_per_pass(n)
    assert c == 0 and g == 64 * live + 2 * 4 * 64 * 4
    mg, mc = bench.diag_lanes_per_pass(n, mix=True)
    assert mg + mc == g
  g, c = bench.diag_lanes_per_pass(1000, mix=True)
  assert 0.78 < c / (g + c) < 13 / 16


def test_int_coordinates_only_for_integer_instances_within_the_span():
  """The integer distance form is offered only where it equals the FP64 one: integer coordinates
  (|v| <= 2^30) with each axis spanning at most 2^20; anything else keeps the FP64 form."""
  from tsp_gpu_clean_amd.context import INT_COORD_SPAN, int_coordinates

  ok = np.array([[0.0, 5.0], [INT_COORD_SPAN, 7.0], [3.0, -INT_COORD_SPAN + 7.0]])
  got = int_coordinates(ok)
  assert got.dtype == np.int32 and np.array_equal(got, ok.astype(np.int32))
  far = -(2.0**30)
  assert int_coordinates(np.array([[far, 1.0], [far + 5.0, 2.0]])) is not None
  assert int_coordinates(np.array([[0.0, 0.0], [INT_COORD_SPAN + 1.0, 0.0]])) is None  # span
  assert int_coordinates(np.array([[0.0, 0.0], [0.0, INT_COORD_SPAN + 1.0]])) is None
  assert int_coordinates(np.array([[0.5, 0.0], [1.0, 2.0]])) is None  # fractional
  assert int_coordinates(np.array([[0.0, np.nan], [1.0, 2.0]])) is None
  assert int_coordinates(np.array([[2.0**31, 0.0], [2.0**31 + 1, 0.0]])) is None  # beyond int32
  assert int_coordinates(np.zeros((0, 2))) is None
