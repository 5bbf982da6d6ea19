"""Learned candidate-edge scorer: replaces k-nearest candidate lists in 2-opt with the top-k of
each city's 32 nearest neighbours ranked by a small MLP.

Everything runs on the device through ``libtspmi.so``: the 12-feature builder, the
12 -> 128 -> 128 -> 1 MLP (bf16 weights, fp32 accumulation, MFMA) forward and backward, Adam on
fp32 master weights, and the top-k selection that feeds the packed candidate 2-opt kernel.
Training labels are the edges of locally optimal tours produced by the batched memetic GA itself
(``BatchedMemeticGA``: tournament, PMX, inversion, survival and candidate 2-opt to the local
optimum).  Plain k-NN lists and the full neighbourhood stay available as the fallback and as the
yardstick of the tour gap the learned lists cost.

The reference has no learned component; its candidate lists are k-nearest neighbours
(src/distances/matrix.py:140-211, used by two_opt.py:117-149).
"""

from __future__ import annotations

import os
import time
from dataclasses import dataclass

import numpy as np

from . import _native

H, F, NFEAT = 128, 16, 12
OFF_W1, OFF_B1 = 0, H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1
NEIGHBOURS = 32
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "scorer_weights.npz")


def init_params(seed: int = 0) -> np.ndarray:
  """He-normal weights, zero biases; padding features (12..15) zero."""
  rng = np.random.default_rng(seed)
  th = np.zeros(NPARAM, dtype=np.float32)
  th[OFF_W1:OFF_B1].reshape(H, F)[:, :NFEAT] = rng.normal(0.0, np.sqrt(2.0 / NFEAT), (H, NFEAT))
  th[OFF_W2:OFF_B2] = rng.normal(0.0, np.sqrt(2.0 / H), H * H)
  th[OFF_W3:OFF_B3] = rng.normal(0.0, np.sqrt(1.0 / H), H)
  return th


class LearnedCandidateScorer:
  def __init__(self, theta: np.ndarray | None = None, seed: int = 0) -> None:
    torch = _native.require_device()
    if _native.scorer_param_count() != NPARAM:
      raise RuntimeError("libtspmi.so parameter layout differs from the Python scorer")
    th = init_params(seed) if theta is None else np.asarray(theta, dtype=np.float32)
    self.theta = torch.from_numpy(th.copy()).cuda()
    self.m = torch.zeros(NPARAM, dtype=torch.float32, device="cuda")
    self.v = torch.zeros(NPARAM, dtype=torch.float32, device="cuda")
    self.wbf = torch.empty(NPARAM, dtype=torch.int16, device="cuda")
    _native.to_bf16(self.theta, self.wbf)
    self.step = 0
    self.history: list[float] = []

  # ---- inference --------------------------------------------------------------------------------
  def features(self, context):
    torch = _native.require_device()
    n = context.dimension
    X = torch.empty((n * NEIGHBOURS, F), dtype=torch.int16, device="cuda")
    _native.scorer_features(context.get_gpu_coordinates(), context.get_gpu_candidate_list(NEIGHBOURS),
                            context.get_gpu_distances(), X)
    return X

  def scores(self, X):
    torch = _native.require_device()
    z = torch.empty(X.shape[0], dtype=torch.float32, device="cuda")
    _native.mlp_forward(X, self.wbf, z)
    return z

  def candidates(self, context, k: int = 8):
    """(n, k) learned candidate table: the k best-scored of the 32 nearest neighbours."""
    torch = _native.require_device()
    n = context.dimension
    z = self.scores(self.features(context))
    out = torch.empty((n, k), dtype=torch.int32, device="cuda")
    _native.scorer_topk(z, context.get_gpu_candidate_list(NEIGHBOURS), k, out)
    return out

  # ---- training ---------------------------------------------------------------------------------
  def train(self, X, y, steps: int = 300, batch: int = 65536, lr: float = 2e-3, seed: int = 0,
            weight_decay: float = 0.0, log_every: int = 50, lr_final: float | None = None) -> list[float]:
    """Minibatch Adam on (X, y); minibatch rows drawn on the host with a seeded generator.

    The learning rate follows a cosine from ``lr`` to ``lr_final`` (constant when None).
    """
    torch = _native.require_device()
    S = int(X.shape[0])
    batch = min(batch, S - S % 8)
    rng = np.random.default_rng(seed)
    w_elems, p_floats = _native.mlp_work_sizes(batch)
    work = torch.empty(w_elems, dtype=torch.int16, device="cuda")
    partials = torch.empty(p_floats, dtype=torch.float32, device="cuda")
    grad = torch.empty(NPARAM + 1, dtype=torch.float32, device="cuda")
    Xb = torch.empty((batch, F), dtype=torch.int16, device="cuda")
    yb = torch.empty(batch, dtype=torch.float32, device="cuda")
    X32 = X.view(torch.int32).view(S, F // 2)
    y32 = y.view(torch.int32).view(S, 1)
    losses = []
    for it in range(steps):
      idx = torch.from_numpy(rng.integers(0, S, batch, dtype=np.int64).astype(np.int32)).cuda()
      _native.gather_rows(X32, None, None, None, idx, Xb.view(torch.int32).view(batch, F // 2), None)
      _native.gather_rows(y32, None, None, None, idx, yb.view(torch.int32).view(batch, 1), None)
      _native.mlp_grad(Xb, yb, self.wbf, work, partials, grad)
      self.step += 1
      lr_t = lr if lr_final is None else lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * it / max(1, steps - 1)))
      _native.adam(self.theta, self.m, self.v, grad, self.step, lr_t, 0.9, 0.999, 1e-8, weight_decay, self.wbf)
      if log_every and (it % log_every == 0 or it == steps - 1):
        losses.append(float(grad[NPARAM].item()))
    self.history.extend(losses)
    return losses

  def state(self) -> np.ndarray:
    return self.theta.cpu().numpy()

  def save(self, path: str, info: dict | None = None) -> None:
    import json

    np.savez(path, theta=self.state(), step=self.step, info=np.array(json.dumps(info or {})))

  @classmethod
  def load(cls, path: str = DEFAULT_WEIGHTS) -> "LearnedCandidateScorer":
    """Weights saved by ``save`` (default: the shipped scorer trained on 2000 generated instances)."""
    import json

    data = np.load(path, allow_pickle=False)
    s = cls(theta=data["theta"])
    s.step = int(data["step"])
    s.info = json.loads(str(data["info"])) if "info" in data.files else {}
    return s


# ---- training data from the GA ------------------------------------------------------------------


def training_instances(count: int = 2000, seed: int = 11, lo: int = 200, hi: int = 5000):
  """Generated instances: n uniform in [lo, hi]; even ones uniform in [0,1e6]^2, odd ones 20
  Gaussian clusters (sigma 2 % of the box) clipped to the box; integer coordinates."""
  rng = np.random.default_rng(seed)
  out = []
  for i in range(count):
    n = int(rng.integers(lo, hi + 1))
    if i % 2 == 0:
      xy = rng.integers(0, 1_000_001, (n, 2)).astype(np.float64)
    else:
      centres = rng.uniform(0, 1e6, (20, 2))
      xy = np.rint(np.clip(centres[rng.integers(0, 20, n)] + rng.normal(0, 0.02e6, (n, 2)), 0, 1e6))
    out.append(xy)
  return out


@dataclass
class TrainingSet:
  X: object  # (S, 16) int16 (bf16 bits)
  y: object  # (S,) float32
  instances: int
  positives: int
  seconds: float


def build_training_set(instances, population: int = 8, generations: int = 3, chunk: int = 64,
                       seed: int = 5, progress=None) -> TrainingSet:
  """Run the batched memetic GA (candidate 2-opt to the local optimum on every survivor) on each
  chunk of instances and label every (city, 32-NN) pair by the best final tour's edges."""
  torch = _native.require_device()
  from .batch import BatchedMemeticGA

  t0 = time.perf_counter()
  total = sum(NEIGHBOURS * int(xy.shape[0]) for xy in instances)
  X = torch.empty((total, F), dtype=torch.int16, device="cuda")
  y = torch.empty(total, dtype=torch.float32, device="cuda")
  row = 0
  for c0 in range(0, len(instances), chunk):
    part = instances[c0 : c0 + chunk]
    ga = BatchedMemeticGA(part, population_size=population, mutation_rate=0.05, tournament_k=5,
                          elite_count=population, ls_max_iterations=-1, ls_candidate_k=NEIGHBOURS,
                          seeds=[seed * 100003 + c0 + i for i in range(len(part))])
    ga.init_random()
    d = ga.draw(generations)
    best = torch.empty((generations, len(part)), dtype=torch.int32, device="cuda")
    ga.run(d, generations, best)
    best_h = best[-1].cpu().numpy()
    work = torch.empty(max(int(x.shape[0]) for x in part), dtype=torch.int32, device="cuda")
    for i, xy in enumerate(part):
      n = int(xy.shape[0])
      o_xy = int(ga.desc_host[i, 2])
      xy_i = ga.xy[o_xy : o_xy + 2 * n].view(n, 2)
      o_c = int(ga.desc_host[i, 5])
      nn = ga.cand[o_c : o_c + n * NEIGHBOURS].view(n, NEIGHBOURS)
      Xi = X[row : row + n * NEIGHBOURS]
      _native.scorer_features(xy_i, nn, ga.distances(i), Xi)
      tour = ga.population(i)[int(best_h[i])].contiguous()
      _native.scorer_labels(tour, nn, work[:n], y[row : row + n * NEIGHBOURS])
      row += n * NEIGHBOURS
    torch.cuda.synchronize()
    if progress is not None:
      progress(c0 + len(part), time.perf_counter() - t0)
    del ga
  positives = int(y.sum().item())
  return TrainingSet(X, y, len(instances), positives, time.perf_counter() - t0)


def train_on_generated(count: int = 2000, steps: int = 3000, batch: int = 65536, lr: float = 2e-3, seed: int = 0,
                       population: int = 8, generations: int = 3, chunk: int = 64, progress=None):
  """Generate ``count`` instances, label them with GA + 2-opt tours and train a scorer on them.

  Returns (scorer, info) where info records the data-generation and training times and losses.
  """
  torch = _native.require_device()
  inst = training_instances(count=count, seed=11 + seed)
  ts = build_training_set(inst, population=population, generations=generations, chunk=chunk, seed=5 + seed,
                          progress=progress)
  sc = LearnedCandidateScorer(seed=seed)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  losses = sc.train(ts.X, ts.y, steps=steps, batch=batch, lr=lr, seed=seed, log_every=max(1, steps // 20),
                    lr_final=lr / 20)
  torch.cuda.synchronize()
  info = {"instances": count, "samples": int(ts.X.shape[0]), "positives": ts.positives,
          "data_seconds": round(ts.seconds, 2), "train_seconds": round(time.perf_counter() - t0, 2),
          "steps": steps, "batch": batch, "lr": lr, "losses": [round(x, 5) for x in losses]}
  return sc, info


# ---- evaluation: time to a 2 % gap against full-neighbourhood 2-opt -----------------------------


def uniform_coords(n: int, seed: int) -> np.ndarray:
  rng = np.random.default_rng(seed)
  return rng.integers(0, 1_000_001, (n, 2)).astype(np.float64)


def greedy_start(D, P: int, seed: int):
  """P greedy nearest-neighbour tours from distinct random start cities (P <= n)."""
  torch = _native.require_device()
  n = int(D.shape[0])
  starts = np.random.default_rng(seed).choice(n, size=P, replace=P > n).astype(np.int32)
  out = torch.empty((P, n), dtype=torch.int32, device="cuda")
  _native.nn_greedy_init(torch.from_numpy(starts).cuda(), D, out)
  return out


def ls_trace(tours0, D, table=None, k: int = 0, chunk: int = 50, limit_s: float = 600.0, on_chunk=None, xy=None):
  """2-opt (best improvement) to the local optimum in chunks of ``chunk`` iterations.

  ``table`` None runs the full neighbourhood.  With ``xy`` (EUC_2D coordinates) the candidate
  scan computes its distances from the coordinates (same tours, faster once the matrix no longer
  fits the caches).  Returns (trace, tours) with trace a list of (seconds, best cost, mean cost)
  after each chunk, starting at (0, start costs).
  """
  torch = _native.require_device()
  tours = tours0.clone()
  P = int(tours.shape[0])
  cost = torch.empty(P, dtype=torch.float64, device="cuda")
  stats = torch.zeros(3, dtype=torch.int64, device="cuda")
  _native.tour_cost(tours, D, cost)
  trace = [(0.0, float(cost.min().item()), float(cost.mean().item()))]
  torch.cuda.synchronize()
  t = 0.0
  while t < limit_s:
    moves0 = int(stats[1].item())
    t0 = time.perf_counter()
    if table is None:
      _native.two_opt(tours, D, None, 0, chunk, stats)
    elif xy is not None:
      _native.two_opt_packed_euc2d(tours, xy, table, k, chunk, stats)
    else:
      _native.two_opt_packed(tours, D, table, k, chunk, stats)
    _native.tour_cost(tours, D, cost)
    best, mean, moves = cost.min().item(), cost.mean().item(), int(stats[1].item())
    t += time.perf_counter() - t0
    trace.append((t, float(best), float(mean)))
    if on_chunk is not None:
      on_chunk(trace[-1], moves)
    if moves == moves0:
      break
  return trace, tours


def renumber_seconds(tours0, xy, table, k: int, ws, reps: int = 3) -> float:
  """Device seconds a renumbered candidate scan spends before its first pass (the Hilbert order,
  the permuted table and coordinates, the tours renumbered in and out): a one-pass call with the
  workspace less the same call without it, best of ``reps`` each, at least 0."""
  torch = _native.require_device()

  def one(w) -> float:
    best = float("inf")
    for _ in range(reps):
      t = tours0.clone()
      e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
      e0.record()
      _native.two_opt_packed_euc2d(t, xy, table, k, 1, None, ws=w)
      e1.record()
      torch.cuda.synchronize()
      best = min(best, e0.elapsed_time(e1) * 1e-3)
    return best

  return max(0.0, one(ws) - one(None))


def ls_logged(tours0, xy, table, k: int, log_moves: int = 16384, ws=None):
  """Candidate 2-opt (best improvement, distances from coordinates) to the local optimum in one
  launch with a per-move progress log.  Returns (log, tours, launch seconds, clock Hz): log[p, m]
  = {cost, device clock} of tour p at its start (m = 0) and after its m-th move (-1 past its
  last move).  With ``ws`` the scan runs on the renumbered instance (the log's clock starts with
  the scan; :func:`renumber_seconds` gives what comes before it)."""
  torch = _native.require_device()
  tours = tours0.clone()
  P = int(tours.shape[0])
  log = torch.full((P, log_moves + 1, 2), -1, dtype=torch.int64, device="cuda")
  stats = torch.zeros(3, dtype=torch.int64, device="cuda")
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  _native.two_opt_packed_euc2d_log(tours, xy, table, k, -1, log, stats, ws=ws)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  host = log.cpu().numpy()
  moves = int(stats[1].item())
  logged = int((host[:, 1:, 0] >= 0).sum())
  if logged != moves:
    raise RuntimeError(f"pass log holds {logged} of {moves} moves: raise log_moves")
  return host, tours, wall, 1e3 * _native.wall_clock_khz()


def log_time_to(log, hz: float, target: float):
  """Seconds from the first tour's start until the first tour whose cost is <= target (the moment
  the population's best reaches the target), at the resolution of single moves; None if never."""
  cost, clock = log[:, :, 0], log[:, :, 1]
  start = clock[:, 0].min()
  hit = (cost >= 0) & (cost <= target)
  if not hit.any():
    return None
  first = np.where(hit.any(axis=1), np.argmax(hit, axis=1), -1)
  rows = np.nonzero(first >= 0)[0]
  return float((clock[rows, first[rows]].min() - start) / hz)


def time_to(trace, target: float, offset: float = 0.0):
  for t, best, _ in trace:
    if best <= target:
      return t + offset
  return None


def candidate_tables(D, xy, scorer, k_learned: int = 8, knn_ks=(8, 20)):
  """Device candidate tables with their build times: {'knn{k}': (table, k, s), 'learned{k}': ...}."""
  torch = _native.require_device()
  n = int(D.shape[0])
  out = {}
  for k in knn_ks:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cand = torch.empty((n, k), dtype=torch.int32, device="cuda")
    _native.knn(D, k, cand)
    tab = torch.empty((n, _native.candidate_lanes(k), 2), dtype=torch.int32, device="cuda")
    _native.candidate_table(cand, D, tab)
    torch.cuda.synchronize()
    out[f"knn{k}"] = (tab, k, time.perf_counter() - t0)
  if scorer is not None:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nn = torch.empty((n, NEIGHBOURS), dtype=torch.int32, device="cuda")
    _native.knn(D, NEIGHBOURS, nn)
    X = torch.empty((n * NEIGHBOURS, F), dtype=torch.int16, device="cuda")
    _native.scorer_features(xy, nn, D, X)
    z = scorer.scores(X)
    cand = torch.empty((n, k_learned), dtype=torch.int32, device="cuda")
    _native.scorer_topk(z, nn, k_learned, cand)
    tab = torch.empty((n, _native.candidate_lanes(k_learned), 2), dtype=torch.int32, device="cuda")
    _native.candidate_table(cand, D, tab)
    torch.cuda.synchronize()
    out[f"learned{k_learned}"] = (tab, k_learned, time.perf_counter() - t0)
  return out


def evaluate_gap(n: int = 10000, P: int = 256, scorer=None, k_learned: int = 8, gap: float = 0.02, seed: int = 2024,
                 full_chunk: int = 200, cand_chunk: int = 50, full_limit_s: float = 600.0, progress=None,
                 coords: bool | None = None) -> dict:
  """Time for each candidate scheme to bring the population's best tour within ``gap`` of the
  full-neighbourhood 2-opt local optimum reached from the same start tours.

  Every candidate scheme runs the same kernel: the coordinate form of the candidate scan when
  ``coords`` (default: when the matrix outgrows the caches, n > 4096), else the matrix form."""
  from .context import coordinate_candidates_pay, int_coordinates
  torch = _native.require_device()
  xy_host = uniform_coords(n, seed)
  xy = torch.from_numpy(xy_host).cuda()
  ixy = int_coordinates(xy_host)
  # the candidate scans' coordinates: int32 for the integer distance form when the instance allows
  scan_xy = torch.from_numpy(ixy).cuda() if ixy is not None else xy
  D = torch.empty((n, n), dtype=torch.int32, device="cuda")
  _native.dist_euc2d(xy, D)
  start = greedy_start(D, P, seed + 1)
  tables = candidate_tables(D, xy, scorer, k_learned)
  coords = coordinate_candidates_pay(n) if coords is None else bool(coords)
  res = {"n": n, "P": P, "gap": gap, "methods": {},
         "candidate_kernel": ("coordinates, " + ("integer form" if ixy is not None else "FP64 form")) if coords
                             else "matrix"}
  tick = None
  if progress:
    tick = lambda p, m: progress("full-chunk", {"seconds": round(p[0], 2), "best": p[1], "moves": m})  # noqa: E731
  full_trace, _ = ls_trace(start, D, None, 0, full_chunk, full_limit_s, tick)
  c_full = full_trace[-1][1]
  target = (1.0 + gap) * c_full
  res["full"] = {"best": c_full, "mean": full_trace[-1][2], "seconds_to_local_optimum": round(full_trace[-1][0], 4),
                 "seconds_to_gap": time_to(full_trace, target), "start_best": full_trace[0][1],
                 "chunks": len(full_trace) - 1}
  if progress:
    progress("full", res["full"])
  for name, (tab, k, prep) in tables.items():
    if coords:
      # one launch to the local optimum, the population's best known after every move; integer
      # instances on the renumbered instance, its renumbering counted with the candidate set-up
      ws = None
      if scan_xy.dtype == torch.int32:
        ws = torch.empty(_native.packed_euc2d_ws_bytes(n, k), dtype=torch.uint8, device="cuda")
        prep += renumber_seconds(start, scan_xy, tab, k, ws)
      log, tours, wall, hz = ls_logged(start, scan_xy, tab, k, ws=ws)
      cost = torch.empty(P, dtype=torch.float64, device="cuda")
      _native.tour_cost(tours, D, cost)
      final = cost.cpu().numpy()
      best, mean = float(final.min()), float(final.mean())
      t_cross = log_time_to(log, hz, target)
      t_gap = None if t_cross is None else t_cross + prep
      t_opt = wall + prep
      resolution = "one move (device clock in the launch)"
    else:
      trace, _ = ls_trace(start, D, tab, k, cand_chunk)
      best, mean = trace[-1][1], trace[-1][2]
      t_gap = time_to(trace, target, prep)
      t_opt = trace[-1][0] + prep
      resolution = f"one chunk of {cand_chunk} passes (host clock)"
    res["methods"][name] = {"k": k, "prep_seconds": round(prep, 5), "best": best, "mean": mean,
                            "gap_at_optimum": round(best / c_full - 1.0, 5),
                            "seconds_to_gap": None if t_gap is None else round(t_gap, 5),
                            "seconds_to_local_optimum": round(t_opt, 5), "time_resolution": resolution}
    if progress:
      progress(name, res["methods"][name])
  return res
