"""Repeated runs of a configured GA and the reference's summary record for them.

Mirrors the TSP path of the reference benchmark runner (src/benchmarking/runner/
algorithm_runner.py): algorithms are built from the same JSON configuration entries
(``{"metaheuristic": "GeneticAlgorithm", "params": ..., "strategies": {slot: {"class", "params"}}}``
plus ``default_params``, tests/fixtures/configs/algorithms.json), each repetition runs on a fresh
``ProblemContext`` seeded ``seed_base + r`` after one warm-up on an isolated instance
(:516-533, :553-575), and ``summary_payload`` returns the record of ``_build_summary_payload``
(:193-442) with the same keys and statistics (sample standard deviation with Bessel's correction,
gaps in percent of the known optimum, phase-time shares).  VRP fields, database storage and
incremental checkpoint resumption stay with the reference.
"""

from __future__ import annotations

import inspect
import json
import math
import time
from collections.abc import Mapping
from typing import Any

import numpy as np

from .context import ProblemContext
from .ga import GeneticAlgorithm
from .pipeline_steps import PIPELINE_STEP_REGISTRY
from .strategies import STRATEGY_REGISTRY

STRATEGY_SLOTS = ("initialization", "selection", "crossover", "mutation", "local_search", "evaluation")


def adaptive_patience(n: int) -> int:
  """Stagnation window 2 sqrt(n) (algorithm_runner.py:57-70)."""
  return int(2 * math.sqrt(n))


def compute_series_stats(arr: np.ndarray, total_successful: int) -> tuple[float, float, float, float]:
  """(mean, std with ddof=1 when >= 2 values, min, max) over the non-NaN entries (:75-91)."""
  arr = np.asarray(arr, dtype=np.float64)
  if total_successful <= 0 or arr.size == 0:
    return 0.0, 0.0, 0.0, 0.0
  valid = arr[~np.isnan(arr)]
  if valid.size == 0:
    return 0.0, 0.0, 0.0, 0.0
  std = float(np.std(valid, ddof=1)) if valid.size > 1 else 0.0
  return float(np.mean(valid)), std, float(np.min(valid)), float(np.max(valid))


def _strategy(spec: Mapping[str, Any], seed: Any):
  name = spec["class"]
  cls = STRATEGY_REGISTRY.get(name)
  if cls is None:
    raise KeyError(f"strategy class {name!r} is not registered (available: {sorted(STRATEGY_REGISTRY)})")
  params = dict(spec.get("params", {}))
  if "seed" not in params and "seed" in inspect.signature(cls).parameters:
    params["seed"] = seed
  return cls(**params)


def build_algorithm(config: Mapping[str, Any], default_params: Mapping[str, Any] | None = None,
                    seed: Any = None) -> GeneticAlgorithm:
  """A ``GeneticAlgorithm`` from one ``algorithm_configs`` entry of the reference's config format."""
  if config.get("metaheuristic", "GeneticAlgorithm") != "GeneticAlgorithm":
    raise ValueError(f"unsupported metaheuristic {config.get('metaheuristic')!r}")
  params = {**dict(default_params or {}), **dict(config.get("params", {}))}
  if seed is not None:
    params["seed"] = seed
  strategies = {slot: _strategy(spec, params.get("seed")) for slot, spec in config.get("strategies", {}).items()}
  missing = [s for s in STRATEGY_SLOTS if s not in strategies and s != "local_search"]
  if missing:
    raise ValueError(f"config lacks strategy slots {missing}")
  pipeline = None
  if "pipeline" in params:
    pipeline = []
    for step in params.pop("pipeline"):
      cls = PIPELINE_STEP_REGISTRY[step["class"]]
      pipeline.append(cls(strategies.get("local_search"), strategies["evaluation"], **dict(step.get("params", {}))))
  return GeneticAlgorithm(**params, **strategies, **({"pipeline": pipeline} if pipeline is not None else {}))


def run_repetitions(algorithm_name: str, factory, problem, max_generations: int, optimal_cost: float | None,
                    repetitions: int, seed_base: int = 42, patience: int | None = None,
                    warmup: bool = True) -> dict[str, Any]:
  """Run ``repetitions`` seeded evolutions of ``factory(seed)`` and return the summary record."""
  n = int(problem.dimension)
  patience = adaptive_patience(n) if patience is None else patience
  if warmup and repetitions > 0:
    factory(seed_base).warmup(ProblemContext(problem, seed=seed_base))
  results: dict[str, list] = {k: [] for k in ("times", "initial_costs", "final_costs", "gaps", "generations",
                                              "stop_reasons", "seeds", "h2d_bytes", "d2h_bytes", "kernel_launches",
                                              "phase_timings", "best_tours")}
  for r in range(repetitions):
    seed = seed_base + r
    results["seeds"].append(seed)
    context = ProblemContext(problem, seed=seed)
    algo = factory(seed)
    t0 = time.perf_counter()
    try:
      tour, stats = algo.evolve(context, list(range(n)), max_generations, optimal_cost=optimal_cost,
                                patience=patience)
      elapsed = time.perf_counter() - t0
    except Exception:  # recorded as a failed repetition, as the reference does (:659-665)
      for k, v in (("times", np.nan), ("initial_costs", np.nan), ("final_costs", np.nan), ("gaps", np.nan),
                   ("generations", 0), ("stop_reasons", "failed"), ("h2d_bytes", 0), ("d2h_bytes", 0),
                   ("kernel_launches", 0), ("phase_timings", {}), ("best_tours", None)):
        results[k].append(v)
      continue
    final = float(stats["best_fitness"])
    results["times"].append(elapsed)
    results["initial_costs"].append(float(stats["initial_fitness"]))
    results["final_costs"].append(final)
    results["gaps"].append((final - optimal_cost) / optimal_cost * 100.0 if optimal_cost else np.nan)
    results["generations"].append(int(stats.get("generations_completed", 0)))
    results["stop_reasons"].append(stats.get("stop_reason", "completed"))
    results["h2d_bytes"].append(stats.get("h2d_bytes", 0))
    results["d2h_bytes"].append(stats.get("d2h_bytes", 0))
    results["kernel_launches"].append(stats.get("kernel_launches", 0))
    results["phase_timings"].append(stats.get("phase_timings", {}))
    results["best_tours"].append(np.asarray(tour).tolist())
  return summary_payload(results, problem, algorithm_name, has_known_optimal=bool(optimal_cost and optimal_cost > 0))


def summary_payload(results: Mapping[str, list], problem, algorithm_name: str,
                    has_known_optimal: bool = True) -> dict[str, Any]:
  """The summary record of algorithm_runner.py:193-442 for TSP (non-VRP) problems."""
  times = np.array(results["times"], dtype=np.float64)
  init = np.array(results["initial_costs"], dtype=np.float64)
  final = np.array(results["final_costs"], dtype=np.float64)
  gaps = np.array(results["gaps"], dtype=np.float64)
  gens = np.array(results["generations"], dtype=np.float64)
  seeds = np.array(results["seeds"])
  ok_init = ~np.isnan(init) & (init > 0) & ~np.isnan(final)
  improvements = np.full_like(init, np.nan)
  np.divide(init - final, init, out=improvements, where=ok_init)
  improvements *= 100.0
  h2d = np.array(results["h2d_bytes"], dtype=np.float64) if results["h2d_bytes"] else np.zeros(len(times))
  d2h = np.array(results["d2h_bytes"], dtype=np.float64) if results["d2h_bytes"] else np.zeros(len(times))
  valid = ~np.isnan(times)
  total = int(valid.sum())
  tours = [results["best_tours"][i] for i in range(len(valid)) if valid[i] and results["best_tours"][i] is not None]
  valid_final = final[valid]
  best_idx = int(np.nanargmin(valid_final)) if total > 0 and bool(np.any(~np.isnan(valid_final))) else 0
  best_tour = tours[best_idx] if tours and best_idx < len(tours) else None

  raw_pt = [results["phase_timings"][i] for i in range(len(valid)) if valid[i] and i < len(results["phase_timings"])]
  mean_pt: dict[str, dict[str, float]] = {}
  share: dict[str, float] = {}
  share_cv: dict[str, float] = {}
  if raw_pt:
    wall_ms = times[valid] * 1000.0
    for phase in sorted({p for run in raw_pt for p in run}):
      tot = [float(run.get(phase, {}).get("total_ms", 0.0)) for run in raw_pt]
      calls = [int(run.get(phase, {}).get("calls", 0)) for run in raw_pt]
      shares = [(t / w * 100.0) if w > 0 else 0.0 for t, w in zip(tot, wall_ms)]
      mt, mc = float(np.mean(tot)), float(np.mean(calls))
      mean_pt[phase] = {"total_ms": mt, "calls": mc, "mean_ms": mt / mc if mc > 0 else 0.0}
      ms = float(np.mean(shares))
      share[phase] = ms
      sd = float(np.std(shares, ddof=1)) if len(shares) > 1 else 0.0
      share_cv[phase] = sd / ms * 100.0 if ms > 1e-9 else 0.0

  runs = len(h2d)
  mean_h2d = float(np.sum(h2d)) / runs if runs else 0.0
  mean_d2h = float(np.sum(d2h)) / runs if runs else 0.0
  mean_k = float(np.sum(results["kernel_launches"])) / runs if runs else 0.0
  mean_time, std_time, min_time, max_time = compute_series_stats(times, total)
  mean_init, std_init, _, _ = compute_series_stats(init, total)
  mean_cost, std_cost, best_cost, worst_cost = compute_series_stats(final, total)
  if has_known_optimal:
    mean_gap, std_gap, best_gap, _ = compute_series_stats(gaps, total)
    raw_gaps = gaps[valid].tolist()
  else:
    mean_gap = std_gap = best_gap = None
    raw_gaps = None
  mean_g, std_g, min_g, max_g = compute_series_stats(gens, total)
  mean_imp, _, _, _ = compute_series_stats(improvements, total)
  return {
    "algorithm": algorithm_name, "repetitions": len(times), "successful_runs": total,
    "mean_time": mean_time, "std_time": std_time, "min_time": min_time, "max_time": max_time,
    "mean_cost": mean_cost, "std_cost": std_cost, "best_cost": best_cost, "worst_cost": worst_cost,
    "mean_initial_cost": mean_init, "std_initial_cost": std_init,
    "mean_gap": mean_gap, "std_gap": std_gap, "best_gap": best_gap, "mean_improvement": mean_imp,
    "mean_generations": mean_g, "std_generations": std_g, "min_generations": min_g, "max_generations": max_g,
    "mean_h2d_mb": mean_h2d / 1e6, "mean_d2h_mb": mean_d2h / 1e6, "total_transfer_mb": (mean_h2d + mean_d2h) / 1e6,
    "mean_kernels": mean_k, "best_tour": best_tour, "raw_best_tours": tours,
    "problem_type": getattr(problem, "problem_type", "TSP"), "edge_type": getattr(problem, "edge_type", "EUC_2D"),
    "raw_times": times[valid].tolist(), "raw_initial_costs": init[valid].tolist(), "raw_costs": final[valid].tolist(),
    "raw_gaps": raw_gaps, "raw_generations": gens[valid].tolist(),
    "raw_stop_reasons": [results["stop_reasons"][i] for i in range(len(valid)) if valid[i]],
    "raw_seeds": seeds[valid].tolist(), "raw_phase_timings": raw_pt, "mean_phase_timings": mean_pt,
    "phase_share_pct": share, "phase_share_cv_pct": share_cv,
  }


def save_record(path: str, record: Mapping[str, Any]) -> None:
  """Write a summary record as JSON (NaN as null)."""

  def clean(v):
    if isinstance(v, float) and math.isnan(v):
      return None
    if isinstance(v, dict):
      return {k: clean(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
      return [clean(x) for x in v]
    if isinstance(v, np.generic):
      return clean(v.item())
    return v

  with open(path, "w", encoding="utf-8") as fh:
    json.dump(clean(dict(record)), fh, indent=1)
