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


def test_batch_graph_replay_and_autotune_keep_state():
  import torch

  from tsp_gpu_clean_amd.batch import BatchedMemeticGA

  say("batch graph replay + autotune")
  coords = _coords([50, 500, 120, 75, 300], seed=2)
  P, G = 32, 4
  ga = BatchedMemeticGA(coords, population_size=P, elite_count=8, ls_candidate_k=20, seeds=[5, 6, 7, 8, 9])
  ga.init_random()
  d = ga.draw(G)
  start = [t.clone() for t in ga.pop], [k.clone() for k in ga.keys]
  shapes = ga.autotune(d, g=0)
  assert set(shapes) == {"gen_block", "gen_grid", "ls_block", "ls_grid"}
  for a, b in zip(ga.pop, start[0]):
    assert torch.equal(a, b)
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
