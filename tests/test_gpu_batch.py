"""Uneven multi-instance batch: every instance evolves exactly as it would alone.

Checked against the single-instance device loop (itself pinned against the oracle GA) and,
for two instances, against the oracle GA directly; plus graph replays and autotuning leaving
the state unchanged.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import ga as oga
from oracle import init as oinit

pytestmark = pytest.mark.gpu


def say(msg):
  print(msg, flush=True)


def _coords(ns, seed=5):
  rng = np.random.default_rng(seed)
  return [rng.integers(0, 1_000_001, (int(n), 2)).astype(np.float64) for n in ns]


@pytest.mark.parametrize("kc,ns", [(8, [6, 300, 57, 120, 9, 211, 64, 33]), (20, [6, 300, 57, 120, 9, 211, 64, 33]),
                                   (8, [4200, 300, 57])])  # the last batch mixes coordinate and matrix scans
def test_batch_matches_single_instance_loops(kc, ns):
  import torch

  import tsp_gpu_clean_amd as t
  from tsp_gpu_clean_amd.batch import BatchedMemeticGA
  from tsp_gpu_clean_amd.engine import DeviceMemeticLoop, draw_generations, upload
  from tsp_gpu_clean_amd.rng import draw_row_seeds

  say(f"batch vs single loops kc={kc} ns={ns}")
  coords = _coords(ns)
  P, G, E = 16, 6, 4
  seeds = [100 + i for i in range(len(ns))]
  ga = BatchedMemeticGA(coords, population_size=P, mutation_rate=0.3, tournament_k=5, elite_count=E,
                        ls_max_iterations=5, ls_candidate_k=kc, seeds=seeds)
  ga.init_random()
  d = ga.draw(G)
  best = torch.empty((G, len(ns)), dtype=torch.int32, device="cuda")
  ga.run(d, G, best)
  for i, (xy, s) in enumerate(zip(coords, seeds)):
    ctx = t.ProblemContext(t.Problem.from_coordinates(f"b{i}", xy), seed=s)
    np.testing.assert_array_equal(ga.distances(i).cpu().numpy(), ctx.get_gpu_distances().cpu().numpy())
    seeds_i = draw_row_seeds(ctx.rng, P)
    pop0 = torch.from_numpy(oinit.random_init(np.arange(ns[i]), P, seeds_i)).cuda()
    draws = upload(draw_generations(ctx.rng, G, P, ns[i], 5))
    loop = DeviceMemeticLoop(ctx, P, 0.3, tournament_k=5, elite_count=E, ls_max_iterations=5, ls_candidate_k=kc)
    loop.load(pop0)
    bi = torch.empty(G, dtype=torch.int32, device="cuda")
    loop.run(draws, G, bi)
    np.testing.assert_array_equal(ga.population(i).cpu().numpy(), loop.population.cpu().numpy(), err_msg=f"inst {i}")
    np.testing.assert_array_equal(ga.fitness(i).cpu().numpy(), loop.fitness.cpu().numpy())
    np.testing.assert_array_equal(best[:, i].cpu().numpy(), bi.cpu().numpy())


def test_batch_matches_oracle_ga():
  import tsp_gpu_clean_amd as t
  from tsp_gpu_clean_amd.batch import BatchedMemeticGA

  say("batch vs oracle GA")
  ns = [40, 90]
  coords = _coords(ns, seed=9)
  P, G = 12, 5
  ga = BatchedMemeticGA(coords, population_size=P, mutation_rate=0.2, elite_count=3, ls_max_iterations=4,
                        ls_candidate_k=6, seeds=[1, 2])
  ga.init_random()
  d = ga.draw(G)
  ga.run(d, G)
  for i, xy in enumerate(coords):
    ctx = t.ProblemContext(t.Problem.from_coordinates("o", xy), seed=i + 1)
    rng = np.random.default_rng(i + 1)
    pop0 = oinit.random_init(np.arange(ns[i]), P, oinit.draw_row_seeds(rng, P))
    res = oga.evolve(rng, ctx.get_cpu_distances(), pop0, max_generations=G, mutation_rate=0.2, tournament_k=5,
                     local_search=oga.LocalSearchSpec("best", 4, ctx.get_cpu_candidate_list(6)), elite_count=3,
                     patience=10**6, record=True)
    np.testing.assert_array_equal(ga.population(i).cpu().numpy(), res.populations[-1])
    np.testing.assert_array_equal(ga.fitness(i).cpu().numpy(), res.fitnesses[-1])


@pytest.mark.parametrize("fused", [False, True])
def test_batch_graph_replay_and_autotune_keep_state(fused):
  import torch

  from tsp_gpu_clean_amd.batch import BatchedMemeticGA

  say(f"batch graph replay + autotune, fused={fused}")
  coords = _coords([50, 500, 120, 75, 300], seed=2)
  P, G = 32, 4
  ga = BatchedMemeticGA(coords, population_size=P, elite_count=8, ls_candidate_k=20, seeds=[5, 6, 7, 8, 9])
  ga.init_random()
  d = ga.draw(G)
  start = [t.clone() for t in ga.pop], [k.clone() for k in ga.keys]
  shapes = ga.autotune(d, g=0)
  shapes["fused"] = fused
  assert set(shapes) == {"fused", "fused_block", "fused_grid", "children_per_task", "gen_block", "gen_grid",
                         "ls_block", "ls_grid"}
  assert set(ga.tuning["fused_generation_ms"]) == {"256x4", "512x4", "256x8", "512x8"}
  assert ga.tuning["staged_generation_ms"] > 0 and ga.children_per_task == shapes["children_per_task"]
  # the state is the current population and its parents' keys (the other buffers are rewritten
  # by every generation before they are read)
  assert torch.equal(ga.pop[ga.cur], start[0][ga.cur])
  for i in range(ga.B):
    o = 2 * i * P
    assert torch.equal(ga.keys[ga.cur][o : o + P], start[1][ga.cur][o : o + P])
  ga.run(d, G)
  want = ga.pop[ga.cur].cpu().numpy(), ga.keys[ga.cur].cpu().numpy()
  for src, dst in zip(start[0], ga.pop):
    dst.copy_(src)
  for src, dst in zip(start[1], ga.keys):
    dst.copy_(src)
  graph = ga.capture(d, G)
  for rep in range(3):
    for src, dst in zip(start[0], ga.pop):
      dst.copy_(src)
    for src, dst in zip(start[1], ga.keys):
      dst.copy_(src)
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ga.pop[ga.cur].cpu().numpy(), want[0], err_msg=f"replay {rep}")
    np.testing.assert_array_equal(ga.keys[ga.cur].cpu().numpy(), want[1])


def _pair(coords, P, E, kc, iters, seeds, rate=0.3):
  """The same batch twice: one run stage by stage, one as a fused launch per generation."""
  from tsp_gpu_clean_amd.batch import BatchedMemeticGA

  out = []
  for fused in (False, True):
    ga = BatchedMemeticGA(coords, population_size=P, mutation_rate=rate, tournament_k=5, elite_count=E,
                          ls_max_iterations=iters, ls_candidate_k=kc, seeds=seeds)
    ga.shapes["fused"] = fused
    ga.shapes["fused_block"] = 512 if kc > 8 else 256
    ga.set_children_per_task(2 if kc > 8 else 4)
    ga.init_random()
    out.append(ga)
  return out


def _same_state(staged, fused, best_s, best_f):
  import torch

  torch.cuda.synchronize()
  np.testing.assert_array_equal(fused.pop[fused.cur].cpu().numpy(), staged.pop[staged.cur].cpu().numpy())
  for i in range(staged.B):  # parents' keys (the offspring halves are scratch)
    np.testing.assert_array_equal(fused.fitness(i).cpu().numpy(), staged.fitness(i).cpu().numpy(), err_msg=f"inst {i}")
  np.testing.assert_array_equal(best_f.cpu().numpy(), best_s.cpu().numpy())
  np.testing.assert_array_equal(fused.stats.cpu().numpy(), staged.stats.cpu().numpy())
  fused.verify()
  c = fused.counters.cpu().numpy()
  B = staged.B
  assert c[3] > 0 and (c[4 + B : 4 + 2 * B] == c[3]).all(), "survivor readiness is not the launch count"
  assert not c[:3].any() and not c[4 : 4 + B].any() and not c[4 + 2 * B :].any(), "a fused launch left a counter set"


@pytest.mark.parametrize("P,E,kc,iters,ns", [
    (16, 4, 8, 5, [6, 300, 57, 120, 9, 211, 64, 33]),
    (18, 5, 20, 10, [4, 700, 5, 1500, 90]),  # a partial copy task, n = 4 and 5
    (15, 0, 8, 5, [30, 200, 64]),  # odd P (a single last child), no elite rows
    (12, 12, 12, 3, [40, 333, 80]),  # every survivor an elite row, no copy task
    (10, 4, 8, 0, [25, 130]),  # elite rows copied and costed without refinement
    (16, 6, 20, 10, [4200, 300, 57]),  # max_n beyond the staged generation's default tiles
])
def test_fused_generation_matches_staged(P, E, kc, iters, ns):
  import torch

  say(f"fused vs staged P={P} E={E} kc={kc} iters={iters} ns={ns}")
  coords = _coords(ns, seed=11)
  G = 6
  staged, fused = _pair(coords, P, E, kc, iters, [40 + i for i in range(len(ns))])
  d = staged.draw(G)
  fused.draw(G)  # keep the two generators in step (draws are per instance and identical)
  best_s = torch.empty((G, len(ns)), dtype=torch.int32, device="cuda")
  best_f = torch.empty_like(best_s)
  staged.run(d, G, best_s)
  fused.run(d, G, best_f)
  _same_state(staged, fused, best_s, best_f)


@pytest.mark.parametrize("fused", [False, True])
def test_batch_integer_distances_match_fp64(fused):
  """The batch on integer instances with every distance from int32 coordinates (integer form)
  evolves exactly as with the FP64 form: populations, keys, best rows and counters, staged and
  fused."""
  import torch

  from tsp_gpu_clean_amd.batch import BatchedMemeticGA

  ns = [4200, 300, 57, 6, 1500, 90]
  say(f"batch integer vs FP64 distances fused={fused} ns={ns}")
  coords = _coords(ns, seed=13)
  G = 5
  runs = []
  for integer in (False, True):
    ga = BatchedMemeticGA(coords, population_size=16, mutation_rate=0.3, tournament_k=5, elite_count=5,
                          ls_max_iterations=10, ls_candidate_k=20, seeds=[70 + i for i in range(len(ns))],
                          integer_distances=integer)
    assert (ga.ixy is not None) == integer
    ga.shapes["fused"] = fused
    ga.shapes["fused_block"] = 512
    ga.init_random()
    d = ga.draw(G)
    best = torch.empty((G, len(ns)), dtype=torch.int32, device="cuda")
    ga.run(d, G, best)
    if fused:
      ga.verify()
    torch.cuda.synchronize()
    runs.append((ga.pop[ga.cur].cpu().numpy(), [ga.fitness(i).cpu().numpy() for i in range(len(ns))],
                 best.cpu().numpy(), ga.stats.cpu().numpy()))
  np.testing.assert_array_equal(runs[0][0], runs[1][0])
  for a, b in zip(runs[0][1], runs[1][1]):
    np.testing.assert_array_equal(a, b)
  np.testing.assert_array_equal(runs[0][2], runs[1][2])
  np.testing.assert_array_equal(runs[0][3], runs[1][3])


def test_fused_generation_config4_full_batch_matches_staged():
  import torch

  from tsp_gpu_clean_amd.batch import c4_instances

  coords = c4_instances(256, seed=7)
  G = 4
  say(f"fused vs staged at config 4: 256 instances, P=64, {G} generations, one eager and one captured run")
  staged, fused = _pair(coords, 64, 16, 20, 10, list(range(256)), rate=0.05)
  d = staged.draw(2 * G)
  best_s = torch.empty((2 * G, 256), dtype=torch.int32, device="cuda")
  best_f = torch.empty_like(best_s)
  staged.run(d, 2 * G, best_s)
  fused.run(d, G, best_f[:G])
  graph = fused.capture(d, G, best_f[G:], start=G)
  graph.replay()
  _same_state(staged, fused, best_s, best_f)
