"""CPU-side checks: the C ABI library loads and exports every declared symbol, the host
logic (registry, pipeline composition, argument validation, fail-loud without a device)."""

from __future__ import annotations

import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tspmi.h")


def _declared_symbols() -> list[str]:
  text = open(HEADER, encoding="utf-8").read()
  return sorted(set(re.findall(r"\b(tspmi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
  import ctypes

  from tsp_gpu_clean_amd import _native

  lib = _native.load_library()
  declared = _declared_symbols()
  assert len(declared) >= 16
  for name in declared:
    assert hasattr(lib, name), f"{name} declared in tspmi.h but not exported"
    assert isinstance(getattr(lib, name), ctypes._CFuncPtr)
  assert set(declared) == set(_native.EXPORTED), "ctypes signature table and header disagree"
  assert lib.tspmi_version() == 1


def test_abi_rejects_bad_arguments_without_touching_a_device():
  from tsp_gpu_clean_amd import _native

  lib = _native.load_library()
  assert lib.tspmi_two_opt(None, 4, 10, None, 10, None, 0, 7, 10, None, None) == _native.TSPMI_EINVAL
  assert b"bad mode" in lib.tspmi_last_error()
  assert lib.tspmi_two_opt(None, 1, 3, None, 3, None, 0, 0, 10, None, None) == 0  # n < 4: no-op
  assert lib.tspmi_rank_select(None, 4, 5, None, None) == _native.TSPMI_EINVAL
  assert lib.tspmi_knn(None, 10, 10, 12, None, None, None) == _native.TSPMI_EINVAL
  assert lib.tspmi_nn_init(None, 1, None, 5, None, 65, None, None, None) == _native.TSPMI_EINVAL
  assert lib.tspmi_dist_euc2d(None, 70000, None, 70000, None, None) == _native.TSPMI_EINVAL


def test_gpu_strategies_fail_loud_without_device(monkeypatch):
  import torch

  import tsp_gpu_clean_amd as t
  from tsp_gpu_clean_amd.errors import GPUUnavailableError

  monkeypatch.setattr(torch.cuda, "is_available", lambda: False)
  prob = t.Problem.from_coordinates("p", np.random.default_rng(0).random((10, 2)) * 100)
  ctx = t.ProblemContext(prob, seed=1)
  routes = np.tile(np.arange(10, dtype=np.int32), (2, 1))
  with pytest.raises(GPUUnavailableError):
    t.LocalSearchGPUTourTwoOpt().improve_batch(routes, ctx)
  with pytest.raises(GPUUnavailableError):
    t.EvaluationGPUCost().evaluate_batch(routes, ctx)
  with pytest.raises(GPUUnavailableError):
    t.SelectionGPUTournament().select_batch(routes, np.zeros(2), 2, ctx)
  with pytest.raises(GPUUnavailableError):
    t.InitializationGPUPermRandom().initialize(list(range(10)), 4, ctx)


def test_constructor_validation():
  import tsp_gpu_clean_amd as t

  with pytest.raises(ValueError):
    t.LocalSearchGPUTourTwoOpt(mode="bogus")
  with pytest.raises(ValueError):
    t.LocalSearchGPUTourTwoOpt(max_iterations=-1)
  with pytest.raises(ValueError):
    t.SelectionGPUTournament(k=0)
  with pytest.raises(ValueError):
    t.InitializationGPUTourNearestNeighbor(k=0)
  with pytest.raises(ValueError):
    t.EliteOnlyLocalSearchStep(t.LocalSearchNoOp(), t.EvaluationGPUCost(), elite_count=0)
  with pytest.raises(ValueError):
    t.PeriodicLocalSearchStep(t.LocalSearchNoOp(), interval=0)
  with pytest.raises(ValueError):
    t.Problem(name="x", dimension=3, coordinates=None, distances=None)


def test_register_strategy_contract():
  import tsp_gpu_clean_amd as t
  from tsp_gpu_clean_amd.strategies import STRATEGY_REGISTRY

  class MyLocalSearch:
    def improve(self, route, context):
      return route

    def improve_batch(self, routes, context):
      return routes

  class NotAStrategy:
    def run(self):
      pass

  t.register_strategy("MyLocalSearchTest", MyLocalSearch)
  assert STRATEGY_REGISTRY["MyLocalSearchTest"] is MyLocalSearch
  with pytest.raises(ValueError):
    t.register_strategy("MyLocalSearchTest", MyLocalSearch)
  t.register_strategy("MyLocalSearchTest", MyLocalSearch, overwrite=True)
  with pytest.raises(TypeError):
    t.register_strategy("Bad", NotAStrategy)
  with pytest.raises(TypeError):
    t.register_strategy("Bad", MyLocalSearch())
  with pytest.raises(ValueError):
    t.register_strategy("  ", MyLocalSearch)
  del STRATEGY_REGISTRY["MyLocalSearchTest"]
  for name, cls in list(STRATEGY_REGISTRY.items()):
    t.register_strategy(name, cls, overwrite=True)  # every built-in satisfies a protocol


def test_pipeline_composition_follows_stage_anchors():
  import tsp_gpu_clean_amd as t
  from tsp_gpu_clean_amd import ga as g

  ls = t.LocalSearchGPUTourTwoOpt()
  ev = t.EvaluationGPUCost()
  kw = dict(initialization=t.InitializationGPUPermRandom(), selection=t.SelectionGPUTournament(),
            crossover=t.CrossoverGPUPermPartiallyMapped(), mutation=t.MutationGPUPermInversion(), evaluation=ev,
            local_search=ls)
  alg = t.GeneticAlgorithm(population_size=8, local_search_stage="none",
                           pipeline=[t.PostInitLocalSearchStep(ls, ev), t.EliteOnlyLocalSearchStep(ls, ev, 4),
                                     t.PeriodicLocalSearchStep(ls, 5, "offspring")], **kw)
  kinds = [type(s).__name__ for s in alg.pipeline]
  assert kinds == ["PostInitLocalSearchStep", "SelectionStep", "CrossoverStep", "MutationStep",
                   "PeriodicLocalSearchStep", "EvaluationStep", "SurvivalStep", "EliteOnlyLocalSearchStep"]
  for stage, expect in [("post_mutation", "LocalSearchStep"), ("post_crossover", "LocalSearchStep")]:
    a = t.GeneticAlgorithm(population_size=8, local_search_stage=stage, **kw)
    assert expect in [type(s).__name__ for s in a.pipeline]
  assert g.STAGE_SPECS["post_evolution"].post_loop_hook
  with pytest.raises(ValueError):
    t.GeneticAlgorithm(population_size=8, local_search_stage="bogus", **kw)
  assert g.GeneticAlgorithm._candidate_beats_incumbent(10.0, True, 10.0 + 2e-6, True)
  assert not g.GeneticAlgorithm._candidate_beats_incumbent(10.0, True, 10.0 + 5e-7, True)
  assert g.GeneticAlgorithm._candidate_beats_incumbent(10.0, True, 10.0, True, allow_equal=True)


def test_host_rng_draw_helpers_match_reference_order():
  from tsp_gpu_clean_amd.rng import draw_row_seeds, make_rng, resolve_rng, sample_distinct_pairs

  from oracle import init as oinit
  from oracle import operators as ops

  a = sample_distinct_pairs(make_rng(3), 50, 0, 9)
  b = ops.sample_distinct_pairs(np.random.default_rng(3), 50, 0, 9)
  assert np.array_equal(a, b)
  assert np.array_equal(draw_row_seeds(make_rng(4), 7), oinit.draw_row_seeds(np.random.default_rng(4), 7))

  class Ctx:
    rng = np.random.default_rng(0)

  assert resolve_rng(Ctx(), 5) is Ctx.rng
  assert isinstance(make_rng([1, 2]), np.random.Generator)


def test_tsplib_reader(berlin52_coords, tmp_path):
  import tsp_gpu_clean_amd as t

  path = tmp_path / "b.tsp"
  lines = ["NAME: b52", "TYPE: TSP", "DIMENSION: 52", "EDGE_WEIGHT_TYPE: EUC_2D", "NODE_COORD_SECTION"]
  lines += [f"{i + 1} {x} {y}" for i, (x, y) in enumerate(berlin52_coords)] + ["EOF"]
  path.write_text("\n".join(lines))
  p = t.load_tsplib_file(str(path))
  assert p.dimension == 52 and p.name == "b52"
  assert np.array_equal(p.coordinates, berlin52_coords)


def test_tsplib_reader_reference_fixtures(berlin52_coords):
  """The reference's fixture instances: GEO, ATT, EUC_2D coordinates and two explicit formats."""
  import tsp_gpu_clean_amd as t
  from tests.conftest import GOLDEN

  kinds = {"burma14": ("GEO", 14), "att48": ("ATT", 48), "berlin52": ("EUC_2D", 52), "gr17": ("EXPLICIT", 17),
           "bayg29": ("EXPLICIT", 29)}
  for name, (edge, n) in kinds.items():
    p = t.load_tsplib_file(os.path.join(GOLDEN, f"{name}.tsp"))
    assert (p.edge_type, p.dimension, p.name) == (edge, n, name)
    if edge == "EXPLICIT":
      d = p.distances
      assert d.shape == (n, n) and np.array_equal(d, d.T) and (np.diag(d) == 0).all() and (d[~np.eye(n, dtype=bool)] > 0).all()
  np.testing.assert_array_equal(t.load_tsplib_file(os.path.join(GOLDEN, "berlin52.tsp")).coordinates, berlin52_coords)
  gr17 = t.load_tsplib_file(os.path.join(GOLDEN, "gr17.tsp")).distances  # LOWER_DIAG_ROW: "0 / 633 0 / 257 390 0 / 91 661 228 0"
  assert gr17[1, 0] == 633 and gr17[2, :2].tolist() == [257, 390] and gr17[3, :3].tolist() == [91, 661, 228]
  bayg = t.load_tsplib_file(os.path.join(GOLDEN, "bayg29.tsp")).distances  # UPPER_ROW: "97 205 139 ..." then "129 103 ..."
  assert bayg[0, 1:4].tolist() == [97, 205, 139] and bayg[1, 2:4].tolist() == [129, 103]


def test_explicit_matrix_formats_agree():
  from tsp_gpu_clean_amd.problem import explicit_matrix

  rng = np.random.default_rng(4)
  n = 7
  a = rng.integers(1, 100, (n, n))
  full = np.triu(a, 1) + np.triu(a, 1).T
  iu, il = np.triu_indices(n, 1), np.tril_indices(n, -1)
  iud, ild = np.triu_indices(n, 0), np.tril_indices(n, 0)
  cases = {"FULL_MATRIX": full.ravel(), "UPPER_ROW": full[iu], "LOWER_ROW": full[il], "UPPER_DIAG_ROW": full[iud],
           "LOWER_DIAG_ROW": full[ild], "UPPER_COL": full.T[il], "LOWER_COL": full.T[iu],
           "UPPER_DIAG_COL": full.T[ild], "LOWER_DIAG_COL": full.T[iud]}
  for fmt, vals in cases.items():
    np.testing.assert_array_equal(explicit_matrix(vals, n, fmt), full, err_msg=fmt)
  with pytest.raises(ValueError):
    explicit_matrix(full[iu][:-1], n, "UPPER_ROW")
