"""Benchmark of the memetic-TSP hot path on one MI355X.

Headline (value): population-batched 2-opt deltas/s on n=1000 integer coordinates uniform in
[0, 10^6]^2 (seed 42), int32 nint distances, P=1024 SplitMix64 random permutations, best
improvement over the full O(n^2) neighbourhood, every tour iterated to its local optimum.
One step = one launch refining all 1024 tours from the same start population.

Also reported on the same line: the same workload with candidate_k=20 lists; GA generations/s of
the fused memetic generation (n=1000 clustered, P=1024, tournament 5, PMX, inversion 0.05,
elite-16 2-opt k=20 max_iter 10) over 200 generations; the uneven multi-instance batch (256
instances, n in [50, 2000], P=64, 100 graph-captured generations); and the learned candidate
scorer's time to a 2 % gap against full-neighbourhood 2-opt at n=10000, P=256 (learned top-8 vs
kNN-20 and kNN-8, the full-neighbourhood yardstick run live).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-scorer-eval]
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
BYTES_PER_DELTA = 8  # two int32 distance gathers per 2-opt delta
CUS = 256  # MI355X compute units
PEAK_CLOCK_HZ = 2.4e9  # MI355X peak engine clock
# Random 4-byte lanes per clock per CU that the texture addresser (TA) takes in when the 64 lanes of
# a gather fall at random inside one 4 KB row: one such wave64 load occupies the TA for 32 cycles
# (TA_TA_BUSY / TA_FLAT_READ_WAVEFRONTS of the scan, and the tspmi_probe_gather micro-benchmark).
GATHER_LANES_PER_CLK_CU = 2.0
GATHER_PEAK = GATHER_LANES_PER_CLK_CU * CUS * PEAK_CLOCK_HZ  # lanes/s
# VALU issue of the headline launch by its counters (SQ_INSTS_VALU x 4 cycles over 1024 SIMDs), and
# the same with every issue cycle counted (SQ_ACTIVE_INST_VALU: the FP32 square root takes two).
VALU_BUSY_BY_COUNTERS = 0.82
VALU_ACTIVE_BY_COUNTERS = 0.90
N_CITIES = 1000
POP = 1024


def c2_inputs(torch, native):
  from tsp_gpu_clean_amd.rng import draw_row_seeds

  xy = np.random.default_rng(42).integers(0, 1_000_001, (N_CITIES, 2)).astype(np.float64)
  D = torch.empty((N_CITIES, N_CITIES), dtype=torch.int32, device="cuda")
  native.dist_euc2d(torch.from_numpy(xy).cuda(), D)
  seeds = draw_row_seeds(np.random.default_rng(42), POP)
  tours = torch.empty((POP, N_CITIES), dtype=torch.int32, device="cuda")
  native.random_init(torch.from_numpy(seeds.view(np.int64)).cuda(),
                     torch.arange(N_CITIES, dtype=torch.int32, device="cuda"), tours)
  return xy, D, tours


MIX_PERIOD, MIX_ROWS = 16, 0b1111101111011110  # 13 rows of each 16 from the coordinates


def diag_lanes_per_pass(n: int, qmax: int = 4, waves: int = 4, mix: bool = False) -> tuple[int, int]:
  """(distance-gather lanes, coordinate lanes) of one pass of the full-neighbourhood steepest scan
  over one tour (256 threads): the carry set-up (two gathers per slot lane), row 0 over every slot,
  then per wave the slot phases -- slot q of wave w runs while i < (n-2) - off(w, q), with
  off(w, q) = 64 (4q + w) on even q and 64 (4q + 3 - w) on odd q.  With `mix`, each phase takes its
  rows in periods of 16 of which all but rows 0, 6 and 11 compute the distances from LDS
  coordinates, and the phase's last (< 16) rows gather.  Not counted: the exact re-walk of the
  pass's candidate diagonals (normally one diagonal, two gathers per row)."""
  rows = n - 2
  gather_rows = coord_rows = 0
  for w in range(waves):
    off = [64 * (waves * q + ((waves - 1 - w) if q & 1 else w)) for q in range(qmax)]
    gather_rows += qmax
    i = 1
    for ql in range(qmax, 0, -1):
      end = rows - off[ql - 1]
      if end > i:
        span = end - i
        if mix:
          periods = span // MIX_PERIOD
          c = periods * bin(MIX_ROWS).count("1")
          coord_rows += ql * c
          gather_rows += ql * (span - c)
        else:
          gather_rows += ql * span
        i = end
  return 64 * gather_rows + 2 * qmax * 64 * waves, 64 * coord_rows


def diag_gather_lanes_per_pass(n: int, qmax: int = 4, waves: int = 4) -> int:
  """Distance-gather lanes of one pass of the matrix-only scan (two_opt_kernel)."""
  return diag_lanes_per_pass(n, qmax, waves)[0]


def gather_probe(torch, native, reps: int = 3) -> dict:
  """Live rate of random 4-byte gathers inside 4 KB rows (tspmi_probe_gather), 4 workgroups per CU."""
  rows = torch.randint(0, 1 << 20, (64, 1024), dtype=torch.int32, device="cuda")
  out = torch.zeros(1, dtype=torch.int32, device="cuda")
  blocks, iters = 4 * CUS, 2048
  native.probe_gather(rows, blocks, 8, out)
  stream = torch.cuda.current_stream()
  best = None
  for _ in range(reps):
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    native.probe_gather(rows, blocks, iters, out)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    best = ms if best is None else min(best, ms)
  lanes = blocks * 256 * iters * 8
  rate = lanes / (best * 1e-3)
  return {"lanes_per_s": rate, "lanes_per_clk_per_cu_at_peak_clock": rate / (CUS * PEAK_CLOCK_HZ), "ms": best,
          "shape": f"{blocks} workgroups x 256 threads x {iters * 8} random 4-byte loads in 4 KB rows"}


def c3_coords():
  rng = np.random.default_rng(42)
  centres = rng.uniform(0, 1e6, (20, 2))
  lab = rng.integers(0, 20, N_CITIES)
  xy = centres[lab] + rng.normal(0, 0.02 * 1e6, (N_CITIES, 2))
  return np.rint(np.clip(xy, 0, 1e6))


def time_two_opt(torch, native, D, start, cand, steps, warmup, xy=None):
  """Returns (seconds per step, deltas per step, moves per step, kernel ms per launch, refined tours,
  passes per step).  With device coordinates `xy`, the candidate scan computes every distance from
  them and the full scan 13 rows in 16, on the instance renumbered in Hilbert order inside the
  timed call (tspmi_two_opt_full_euc2d_ordered)."""
  work = torch.empty_like(start)
  stats = torch.zeros(3, dtype=torch.int64, device="cuda")
  if cand is not None:
    table = torch.empty((cand.shape[0], native.candidate_lanes(cand.shape[1]), 2), dtype=torch.int32, device="cuda")
    native.candidate_table(cand, D, table)
    if xy is not None:
      run = lambda: native.two_opt_packed_euc2d(work, xy, table, cand.shape[1], -1, stats)  # noqa: E731
    else:
      run = lambda: native.two_opt_packed(work, D, table, cand.shape[1], -1, stats)  # noqa: E731
  elif xy is not None:
    ws = torch.empty(native.full_euc2d_ws_bytes(start.shape[1]), dtype=torch.uint8, device="cuda")
    run = lambda: native.two_opt_full_euc2d(work, D, xy, -1, stats, ws=ws)  # noqa: E731
  else:
    run = lambda: native.two_opt(work, D, None, 0, -1, stats)  # noqa: E731
  for _ in range(warmup):
    work.copy_(start)
    run()
  torch.cuda.synchronize()
  stats.zero_()
  stream = torch.cuda.current_stream()
  ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
  t0 = time.perf_counter()
  for i in range(steps):
    work.copy_(start)
    ev[i][0].record(stream)
    run()
    ev[i][1].record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  kernel_ms = sum(a.elapsed_time(b) for a, b in ev) / steps
  s = stats.cpu().tolist()
  return wall / steps, s[0] / steps, s[1] / steps, kernel_ms, work, s[2] / steps


def time_generations(torch, native, generations=200, warmup=5):
  from tsp_gpu_clean_amd import Problem, ProblemContext
  from tsp_gpu_clean_amd.engine import DeviceMemeticLoop, draw_generations, upload
  from tsp_gpu_clean_amd.rng import draw_row_seeds

  ctx = ProblemContext(Problem.from_coordinates("c3", c3_coords()), seed=42)
  seeds = draw_row_seeds(ctx.rng, POP)
  pop0 = torch.empty((POP, N_CITIES), dtype=torch.int32, device="cuda")
  native.random_init(torch.from_numpy(seeds.view(np.int64)).cuda(),
                     torch.arange(N_CITIES, dtype=torch.int32, device="cuda"), pop0)
  draws = upload(draw_generations(ctx.rng, generations + warmup, POP, N_CITIES, 5))
  loop = DeviceMemeticLoop(ctx, POP, 0.05, tournament_k=5, elite_count=16, ls_max_iterations=10, ls_candidate_k=20)
  loop.load(pop0)
  for g in range(warmup):
    loop.step(draws, g)
  torch.cuda.synchronize()
  best = torch.empty(generations, dtype=torch.int32, device="cuda")
  start_pop = loop.population.clone()
  start_fit = loop.fitness.clone()
  generations -= generations % 2
  graph = loop.capture(draws, generations, best, start=warmup)
  loop.load(start_pop, start_fit)
  loop.stats.zero_()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  graph.replay()
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  loop.verify()
  best_cost = float(loop.fitness.min().item())
  elite_deltas = int(loop.stats[0].item()) / generations
  return generations / wall, wall, best_cost, elite_deltas


def time_evolve(generations: int, step_generations: int = 40) -> dict:
  """The same GA workload through the public ``GeneticAlgorithm.evolve`` (strategies, pipeline
  steps, incumbent and stopping rules): on the graph-captured device route, and step by step."""
  import tsp_gpu_clean_amd as t

  ctx = t.ProblemContext(t.Problem.from_coordinates("c3", c3_coords()), seed=42)

  def make(device_loop):
    ls = t.LocalSearchGPUTourTwoOpt(mode="best", max_iterations=10, candidate_k=20)
    ev = t.EvaluationGPUCost(return_gpu=True)
    return t.GeneticAlgorithm(population_size=POP, mutation_rate=0.05, initialization=t.InitializationGPUPermRandom(),
                              selection=t.SelectionGPUTournament(k=5), crossover=t.CrossoverGPUPermPartiallyMapped(),
                              mutation=t.MutationGPUPermInversion(), evaluation=ev, local_search=ls,
                              local_search_stage="none", pipeline=[t.EliteOnlyLocalSearchStep(ls, ev, 16)],
                              enable_phase_timing=False, device_loop=device_loop)

  out = {}
  legs = (("device_loop", True, generations), ("device_loop_1000", True, 1000), ("step_pipeline", False, step_generations))
  for name, dl, gens in legs:
    make(dl).evolve(ctx, list(range(N_CITIES)), 40, patience=10**9)  # warm caches, kernels and capture
    ga = make(dl)
    t0 = time.perf_counter()
    _, stats = ga.evolve(ctx, list(range(N_CITIES)), gens, patience=10**9)
    wall = time.perf_counter() - t0
    out[name] = {"generations": stats["generations_completed"], "wall_s": wall,
                 "gens_per_s": stats["generations_completed"] / wall, "best_cost": stats["best_fitness"]}
  out["chunk"] = make(True).device_loop_chunk
  out["note"] = "wall clock of evolve() incl. initialisation, initial evaluation and host draws"
  return out


def time_uneven_batch(torch, generations=100):
  """Config 4: 256 instances, n uniform in [50, 2000], P=64 each, 100 graph-captured generations."""
  from tsp_gpu_clean_amd.batch import BatchedMemeticGA, c4_instances

  coords = c4_instances(256, seed=7)
  t0 = time.perf_counter()
  ga = BatchedMemeticGA(coords, population_size=64, mutation_rate=0.05, tournament_k=5, elite_count=16,
                        ls_max_iterations=10, ls_candidate_k=20, seeds=list(range(256)))
  ga.init_random()
  torch.cuda.synchronize()
  setup = time.perf_counter() - t0
  d = ga.draw(generations + 2)
  shapes = dict(ga.autotune(d, g=generations))
  ga.run(d, 2, start=generations)  # warm-up generations on the spare draw slots
  torch.cuda.synchronize()
  start = [p.clone() for p in ga.pop], [k.clone() for k in ga.keys], ga.cur
  forms, finals = {}, {}
  # both forms from the same state: one persistent launch per generation, and one launch per stage
  for name, fused in (("fused", True), ("staged", False)):
    for src, dst in zip(start[0], ga.pop):
      dst.copy_(src)
    for src, dst in zip(start[1], ga.keys):
      dst.copy_(src)
    ga.cur = start[2]
    ga.shapes["fused"] = fused
    graph = ga.capture(d, generations)
    ga.stats.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    graph.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ga.verify()
    forms[name] = {"wall_s": wall, "gens_per_s": generations / wall, "elite_deltas": int(ga.stats[0].item())}
    finals[name] = ga.pop[ga.cur].clone()
    del graph
  if not torch.equal(finals["fused"], finals["staged"]):
    raise RuntimeError("the fused and staged uneven-batch runs disagree")
  ga.shapes["fused"] = shapes["fused"]
  chosen = forms["fused" if shapes["fused"] else "staged"]
  wall = chosen["wall_s"]
  ns = np.array([c.shape[0] for c in coords])
  gbytes = 16 * 64 * int(ns.sum()) + 8 * chosen["elite_deltas"] / generations
  return {"instances": 256, "n_range": [int(ns.min()), int(ns.max())], "population": 64, "generations": generations,
          "wall_s": wall, "gens_per_s": generations / wall, "instance_gens_per_s": 256 * generations / wall,
          "form": "fused" if shapes["fused"] else "staged", "fused_gens_per_s": forms["fused"]["gens_per_s"],
          "staged_gens_per_s": forms["staged"]["gens_per_s"], "autotune": getattr(ga, "tuning", None),
          "setup_s": setup, "distance_bytes": int(4 * (ns * ns).sum()), "shapes": shapes,
          "bytes_per_generation": gbytes, "achieved_GBps": gbytes * generations / wall / 1e9,
          "hbm_frac": gbytes * generations / wall / 1e9 / HBM_PEAK_GBS,
          "workload": "256 instances, n uniform-integer [50,2000], uniform coords [0,1e6]^2, P=64 each, tournament 5, "
                      "PMX, inversion 0.05, elite-16 2-opt best k=20 max_iter 10, 100 generations captured in one HIP "
                      "graph; form (one persistent launch per generation, or one launch per stage) and launch shapes "
                      "autotuned; both forms timed from the same state, identical results checked"}


C5_SEEDS = (2024, 2025, 2026)


def learned_candidates_eval() -> dict:
  """Config 5: learned top-8 vs kNN-20 (and kNN-8) time to within 2 % of the full 2-opt yardstick,
  on three instance seeds; each candidate scheme's time is read at single-move resolution from
  one logged launch to its local optimum."""
  from tsp_gpu_clean_amd import scorer as S

  sc = S.LearnedCandidateScorer.load()

  def progress(name, r):
    print(f"bench: scorer eval {name} {json.dumps(r)}", file=sys.stderr, flush=True)

  per_seed = {}
  for seed in C5_SEEDS:
    per_seed[seed] = S.evaluate_gap(n=10000, P=256, scorer=sc, k_learned=8, gap=0.02, seed=seed, progress=progress)
  names = ("learned8", "knn8", "knn20")
  times = {m: [per_seed[s]["methods"][m]["seconds_to_gap"] for s in C5_SEEDS] for m in names}

  def stats(v):
    v = [x for x in v if x is not None]
    return {"mean": float(np.mean(v)), "min": float(np.min(v)), "max": float(np.max(v)), "n": len(v)} if v else None

  ratio = lambda a, b: [x / y if (x and y) else None for x, y in zip(times[a], times[b])]  # noqa: E731
  res = per_seed[C5_SEEDS[0]]
  m = res["methods"]
  info = getattr(sc, "info", {})
  return {
    "workload": "n=10000 uniform integer coords [0,1e6]^2 (seeds 2024-2026), P=256 greedy nearest-neighbour starts, "
                "best-improvement 2-opt; yardstick = full-neighbourhood 2-opt local optimum from the same starts; "
                "target = 1.02 x yardstick best; times include each scheme's candidate construction and are read "
                "from the device clock at single-move resolution (one launch to the local optimum per scheme)",
    "full_best": res["full"]["best"], "full_seconds": res["full"]["seconds_to_local_optimum"],
    "full_seconds_to_gap": res["full"]["seconds_to_gap"],
    "learned8_seconds_to_gap": m["learned8"]["seconds_to_gap"], "knn20_seconds_to_gap": m["knn20"]["seconds_to_gap"],
    "knn8_seconds_to_gap": m["knn8"]["seconds_to_gap"],
    "learned8_vs_knn20_speedup": stats(ratio("knn20", "learned8")),
    "width_gain_knn20_over_knn8": stats(ratio("knn20", "knn8")),
    "learning_gain_knn8_over_learned8": stats(ratio("knn8", "learned8")),
    "seconds_to_gap_over_seeds": {k: stats(v) for k, v in times.items()},
    "per_seed": {str(s): {"full_best": per_seed[s]["full"]["best"],
                          **{k: {f: per_seed[s]["methods"][k][f] for f in ("seconds_to_gap", "best", "mean",
                                                                          "gap_at_optimum", "seconds_to_local_optimum")}
                             for k in names}} for s in C5_SEEDS},
    "methods": m,
    "training": {k: info.get(k) for k in ("instances", "samples", "positives", "data_seconds", "train_seconds",
                                          "steps", "batch")},
    "mlp": mlp_throughput(sc),
  }


MLP_FLOP_FWD = 2 * (12 * 128 + 128 * 128 + 128)  # per (city, neighbour) candidate
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA


def mlp_throughput(sc, reps: int = 20) -> dict:
  """Forward over the 320,000 candidates of an n=10000 instance, and one training step (forward,
  backward, weight gradients, reduction, Adam) on 65,536 samples; random bf16 features."""
  import torch

  from tsp_gpu_clean_amd import _native
  from tsp_gpu_clean_amd import scorer as S

  g = torch.Generator(device="cpu").manual_seed(0)
  out = {}
  for S_, kind in ((320_000, "forward"), (65_536, "train_step")):
    x = torch.randn((S_, 16), generator=g)
    x[:, 12:] = 0
    X = x.to(torch.bfloat16).view(torch.int16).cuda()
    stream = torch.cuda.current_stream()
    if kind == "forward":
      z = torch.empty(S_, dtype=torch.float32, device="cuda")
      fn = lambda: _native.mlp_forward(X, sc.wbf, z)  # noqa: E731
      flop = S_ * MLP_FLOP_FWD
    else:
      y = (torch.rand(S_, generator=g) < 0.06).float().cuda()
      w, p = _native.mlp_work_sizes(S_)
      work = torch.empty(w, dtype=torch.int16, device="cuda")
      part = torch.empty(p, dtype=torch.float32, device="cuda")
      grad = torch.empty(S.NPARAM + 1, dtype=torch.float32, device="cuda")
      th, m, v, wb = sc.theta.clone(), torch.zeros_like(sc.m), torch.zeros_like(sc.v), sc.wbf.clone()

      def fn():
        _native.mlp_grad(X, y, wb, work, part, grad)
        _native.adam(th, m, v, grad, 1, 1e-3, 0.9, 0.999, 1e-8, 0.0, wb)
      flop = S_ * 3 * MLP_FLOP_FWD
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
      fn()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    tf = flop / (ms * 1e-3) / 1e12
    out[kind] = {"samples": S_, "ms": ms, "TFLOPs": tf, "mfma_frac": tf / BF16_PEAK_TFLOPS,
                 "flop_per_sample": flop // S_}
  return out


_CPU_SCRIPT = r"""
import json, os, sys, time
sys.path.insert(0, ROOT)
import numpy as np
from oracle import two_opt as o2, init as oinit, ga as oga
from oracle.distances import compute_candidate_list, compute_distance_matrix

def c2_tours():
  xy = np.random.default_rng(42).integers(0, 1_000_001, (N, 2)).astype(np.float64)
  d = compute_distance_matrix(xy)
  return d, oinit.random_init(np.arange(N), P, oinit.draw_row_seeds(np.random.default_rng(42), P))

def refine_row(args):
  os.sched_setaffinity(0, {args[2]})
  d, tours = c2_tours()
  c = o2.new_counter()
  out = o2.improve_batch(tours[args[0]:args[1]], d, None, "best", 10**9, c)
  return out, c.deltas, c.passes

if MODE == "single":
  os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
  d, tours = c2_tours()
  c = o2.new_counter()
  t0 = time.perf_counter(); rows = 0; outs = []
  while rows < P and (time.perf_counter() - t0) < SECONDS:
    outs.append(o2.improve_batch(tours[rows:rows + 1], d, None, "best", 10**9, c)); rows += 1
  wall = time.perf_counter() - t0
  np.save(OUT, np.concatenate(outs))
  print(json.dumps({"deltas": c.deltas, "wall": wall, "rows": rows, "passes": c.passes, "first_row": 0}))
elif MODE == "all":
  import multiprocessing as mp
  cores = sorted(os.sched_getaffinity(0))[:WORKERS]
  jobs = [(FIRST + i, FIRST + i + 1, c) for i, c in enumerate(cores)]
  t0 = time.perf_counter()
  with mp.get_context("fork").Pool(len(jobs)) as pool:
    res = pool.map(refine_row, jobs)
  wall = time.perf_counter() - t0
  np.save(OUT, np.concatenate([r[0] for r in res]))
  print(json.dumps({"deltas": int(sum(r[1] for r in res)), "wall": wall, "rows": len(jobs),
                    "passes": int(sum(r[2] for r in res)), "first_row": FIRST, "workers": len(jobs)}))
else:
  os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
  rng = np.random.default_rng(42)
  centres = rng.uniform(0, 1e6, (20, 2))
  lab = rng.integers(0, 20, N)
  xy = np.rint(np.clip(centres[lab] + rng.normal(0, 0.02 * 1e6, (N, 2)), 0, 1e6))
  d = compute_distance_matrix(xy)
  grng = np.random.default_rng(42)
  pop = oga.init_population(grng, d, N, P, "random")
  ls = oga.LocalSearchSpec("best", 10, compute_candidate_list(d, 20))
  t0 = time.perf_counter()
  r = oga.evolve(grng, d, pop, max_generations=GENS, mutation_rate=0.05, tournament_k=5, local_search=ls,
                 elite_count=16, patience=10**9)
  wall = time.perf_counter() - t0
  print(json.dumps({"generations": r.stats["generations_completed"], "wall": wall, "best": r.stats["best_fitness"]}))
"""


def _run_cpu(mode: str, out_path: str, seconds: float = 20.0, first: int = 0, workers: int = 16, gens: int = 10):
  head = (f"ROOT = {ROOT!r}\nN = {N_CITIES}\nP = {POP}\nMODE = {mode!r}\nOUT = {out_path!r}\n"
          f"SECONDS = {seconds}\nFIRST = {first}\nWORKERS = {workers}\nGENS = {gens}\n")
  env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
  out = subprocess.run([sys.executable, "-c", head + _CPU_SCRIPT], capture_output=True, text=True, env=env,
                       timeout=900, check=True)
  return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline(gpu_tours=None, seconds_hint: float = 20.0) -> dict:
  """Reference-faithful NumPy 2-opt on bounded samples of the C2 workload: one pinned core for
  about ``seconds_hint`` s, then one tour per core on up to 16 cores; the refined sample rows are
  compared with the GPU's rows of the same start population.  Also the oracle GA (C3) on one core
  for 10 generations."""
  import tempfile

  cores = len(os.sched_getaffinity(0))
  workers = min(16, cores)
  with tempfile.TemporaryDirectory() as tmp:
    single = _run_cpu("single", os.path.join(tmp, "single.npy"), seconds=seconds_hint)
    allc = _run_cpu("all", os.path.join(tmp, "all.npy"), first=single["rows"], workers=workers)
    ga = _run_cpu("ga", os.path.join(tmp, "ga.npy"), gens=10)
    exact = None
    if gpu_tours is not None:
      got = np.concatenate([np.load(os.path.join(tmp, "single.npy")), np.load(os.path.join(tmp, "all.npy"))])
      rows = np.arange(got.shape[0])
      exact = bool(np.array_equal(got, gpu_tours[rows]))
  cpu = "unknown"
  try:
    with open("/proc/cpuinfo", encoding="utf-8") as fh:
      cpu = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "unknown")
  except OSError:
    pass
  if exact is False:
    raise RuntimeError("CPU baseline rows differ from the GPU's refined rows")
  return {
    "value": single["deltas"] / single["wall"],
    "unit": "deltas/s",
    "cores": 1,
    "kind": "port",
    "sample": (f"{single['rows']} tour(s) of the same workload (n=1000, full neighbourhood, best, to local optimum; "
               f"{single['passes']} passes, {single['deltas']} deltas in {single['wall']:.1f} s) with the NumPy "
               f"restatement of the reference CPU scan, 1 thread pinned to 1 core of {cpu}; host has {cores} usable "
               f"cores"),
    "all_cores": {"value": allc["deltas"] / allc["wall"], "unit": "deltas/s", "cores": allc["workers"],
                  "sample": f"rows {allc['first_row']}..{allc['first_row'] + allc['rows'] - 1}, one tour per pinned "
                            f"core, {allc['deltas']} deltas in {allc['wall']:.1f} s"},
    "bit_exact_rows": (single["rows"] + allc["rows"]) if exact else None,
    "ga_gens_per_s": {"value": ga["generations"] / ga["wall"], "unit": "generations/s", "cores": 1,
                      "sample": f"oracle GA on the C3 workload (P=1024, n=1000 clustered, elite-16 2-opt k=20), "
                                f"{ga['generations']} generations in {ga['wall']:.1f} s"},
  }


def load_traffic():
  path = os.path.join(ROOT, "profiles", "two_opt_c2_traffic.json")
  if not os.path.exists(path):
    return None
  with open(path, encoding="utf-8") as fh:
    return json.load(fh).get("hbm_bytes_per_launch")


def main() -> int:
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=3)
  ap.add_argument("--warmup", type=int, default=1)
  ap.add_argument("--no-cpu-baseline", action="store_true")
  ap.add_argument("--no-extras", action="store_true")
  ap.add_argument("--no-scorer-eval", action="store_true")
  ap.add_argument("--generations", type=int, default=200)
  ap.add_argument("--mlp-only", action="store_true",
                  help="only time the scorer MLP (forward over 320,000 candidates, training steps on 65,536)")
  ap.add_argument("--copy-probe-mib", type=int, default=0,
                  help="only run a plain device-to-device copy of this many MiB (HBM counter calibration)")
  args = ap.parse_args()
  if args.gpus != 1:
    print(f"bench.py measures one MI355X only: --gpus {args.gpus} is not supported (one GPU per process; "
          "tours shard by population with no exchange).", file=sys.stderr)
    return 2

  import torch

  from tsp_gpu_clean_amd import _native

  _native.require_device()
  if args.copy_probe_mib > 0:
    src = torch.empty(args.copy_probe_mib << 20, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
      dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": args.copy_probe_mib << 20, "copies": 3}))
    return 0
  if args.mlp_only:
    from tsp_gpu_clean_amd import scorer as S

    print(json.dumps({"mlp": mlp_throughput(S.LearnedCandidateScorer.load())}))
    return 0
  print(f"bench: C2 setup n={N_CITIES} P={POP}", file=sys.stderr, flush=True)
  xy, D, start = c2_inputs(torch, _native)
  xy_dev = torch.from_numpy(xy).cuda()
  # the kernel LocalSearchGPUTourTwoOpt runs for this shape: EUC_2D, full neighbourhood, n <= 2050
  sec, deltas, moves, kms, refined, passes = time_two_opt(torch, _native, D, start, None, args.steps, args.warmup,
                                                          xy_dev)
  print(f"bench: full neighbourhood {deltas / sec:.3e} deltas/s, {kms:.1f} ms/launch", file=sys.stderr, flush=True)
  value = deltas / sec
  achieved = deltas * BYTES_PER_DELTA / (kms * 1e-3) / 1e9
  traffic = load_traffic()
  glanes, clanes = diag_lanes_per_pass(N_CITIES, mix=True)
  lanes, coord_lanes = passes * glanes, passes * clanes
  probe = gather_probe(torch, _native)
  msec, mdeltas, _, mkms, mtours, _ = time_two_opt(torch, _native, D, start, None, 1, 0)
  if not torch.equal(mtours, refined):
    raise RuntimeError("matrix-only and mixed full scans disagree on the C2 tours")
  line = {
    "metric": "2-opt deltas/s & gens/s (n=1000,pop=1024); time-to-2% gap, learned vs plain kNN",
    "value": value,
    "unit": "deltas/s",
    "n_gpus": 1,
    "steps": args.steps,
    "warmup": args.warmup,
    "ms_per_step": sec * 1e3,
    "higher_is_better": True,
    "scaling": "weak",
    "vs_baseline": None,
    "dtype": "int32",
    "data": "synthetic",
    "config": {
      "workload": "batched 2-opt best-improvement, full O(n^2) neighbourhood, to local optimum",
      "n": N_CITIES, "population": POP, "coords": "uniform integer [0,1e6]^2 seed 42",
      "init": "SplitMix64 random permutations (seeds from default_rng(42))",
      "deltas_per_step": deltas, "moves_per_step": moves,
    },
    "roofline": {
      "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
      "traffic": traffic,
      "traffic_source": "stored calibration, not measured in this run: profiles/two_opt_c2_traffic.json "
                        "(rocprofv3 FETCH_SIZE + WRITE_SIZE of this launch, scaled by a 256 MiB copy)",
      "measured_hbm_GBps": (traffic / (kms * 1e-3) / 1e9) if traffic else None,
      "measured_hbm_frac": (traffic / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
      "note": "achieved counts the algorithmic 8 B per delta; the 4 MB distance matrix stays in the caches, so the "
              "HBM moves only `traffic` bytes per launch and the binding unit is on chip (roofline_onchip)",
      "kernel": "tspmi::two_opt_mix_kernel<4>", "kernel_ms_per_launch": kms,
      "bytes_per_unit": BYTES_PER_DELTA, "units_per_launch": deltas,
    },
    "roofline_onchip": {
      "bound": "VALU issue (coordinate rows), with texture-address gather issue (TA, gathered rows) below it; "
               "the row split balances the two",
      "valu_busy_by_counters": VALU_BUSY_BY_COUNTERS,
      "valu_active_by_counters": VALU_ACTIVE_BY_COUNTERS,
      "valu_busy_source": "stored: rocprofv3 SQ_INSTS_VALU of this launch, profiles/two_opt_c2_counters.md",
      "achieved": lanes / (kms * 1e-3),
      "peak": GATHER_PEAK, "unit": "gather-lanes/s", "frac": lanes / (kms * 1e-3) / GATHER_PEAK,
      "peak_basis": f"{GATHER_LANES_PER_CLK_CU:g} random 4-byte lanes per clock per CU (a wave64 gather at random "
                    f"offsets of one 4 KB row occupies the TA 32 cycles) x {CUS} CUs x {PEAK_CLOCK_HZ / 1e9:g} GHz",
      "lanes_per_launch": lanes, "lanes_per_pass_per_tour": glanes,
      "coordinate_lanes_per_launch": coord_lanes, "coordinate_lanes_per_pass_per_tour": clanes,
      "passes_per_launch": passes, "probe": probe,
      "note": "13 rows in 16 take their distances from LDS coordinates (VALU: a lower bound from the FP32 root "
              "for integer coordinates, then an exact re-walk of the pass's candidate diagonals) and issue no "
              "gather; the TA fraction counts only the gathered rows",
      "counters": "profiles/two_opt_c2_counters.md (TA busy fraction, L1 hit rate, VALU and LDS issue)",
    },
    "matrix_only": {
      "kernel": "tspmi::two_opt_kernel<4>", "kernel_ms_per_launch": mkms, "deltas_per_s": mdeltas / msec,
      "gather_lanes_per_s": passes * diag_gather_lanes_per_pass(N_CITIES) / (mkms * 1e-3),
      "gather_frac": passes * diag_gather_lanes_per_pass(N_CITIES) / (mkms * 1e-3) / GATHER_PEAK,
      "note": "same tours and counters with every distance gathered from the matrix (checked in this run)",
    },
  }
  if not args.no_extras:
    print("bench: candidate_k=20", file=sys.stderr, flush=True)
    from tsp_gpu_clean_amd.context import coordinate_candidates_pay, int_coordinates

    cand = torch.empty((N_CITIES, 20), dtype=torch.int32, device="cuda")
    _native.knn(D, 20, cand)
    # the distance source the local-search strategy picks for this shape: coordinates, as int32 for
    # the integer distance form (integer coordinates within the 2^20 span)
    cxy = None
    if coordinate_candidates_pay(N_CITIES, POP):
      ixy = int_coordinates(xy)
      cxy = torch.from_numpy(ixy).cuda() if ixy is not None else xy_dev
    s20, d20, _, k20ms, _, _ = time_two_opt(torch, _native, D, start, cand, max(1, args.steps), 1, cxy)
    line["candidate_k20"] = {"deltas_per_s": d20 / s20, "ms_per_step": s20 * 1e3, "kernel_ms": k20ms,
                             "deltas_per_step": d20,
                             "distances": "matrix" if cxy is None else
                                          ("coordinates, integer form" if cxy.dtype == torch.int32 else "coordinates")}
    print(f"bench: GA generations ({args.generations})", file=sys.stderr, flush=True)
    gps, gwall, gbest, gdel = time_generations(torch, _native, args.generations)
    line["gens_per_s"] = gps
    gbytes = 16 * POP * N_CITIES + 8 * gdel
    line["ga"] = {"generations": args.generations, "wall_s": gwall, "best_cost": gbest,
                  "bytes_per_generation": gbytes, "elite_deltas_per_generation": gdel,
                  "achieved_GBps": gbytes * gps / 1e9, "hbm_frac": gbytes * gps / 1e9 / HBM_PEAK_GBS,
                  "workload": "n=1000 clustered (20 Gaussian clusters, sigma 2% of 1e6), P=1024, tournament 5, PMX, "
                              "inversion 0.05, elite-16 2-opt best k=20 max_iter 10"}
    print("bench: GA through GeneticAlgorithm.evolve", file=sys.stderr, flush=True)
    line["ga_evolve"] = time_evolve(args.generations)
    print("bench: uneven batch (256 instances)", file=sys.stderr, flush=True)
    line["uneven_batch"] = time_uneven_batch(torch)
    if not args.no_scorer_eval:
      print("bench: learned candidates vs kNN at n=10000", file=sys.stderr, flush=True)
      line["learned_candidates"] = learned_candidates_eval()
  if not args.no_cpu_baseline:
    print("bench: CPU baseline (NumPy restatement, 1 core)", file=sys.stderr, flush=True)
    line["cpu_baseline"] = cpu_baseline(refined.cpu().numpy())
  print(json.dumps(line), flush=True)
  return 0


if __name__ == "__main__":
  sys.exit(main())
