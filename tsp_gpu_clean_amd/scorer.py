"""Learned candidate-edge scorer: replaces k-nearest candidate lists in 2-opt with the top-k of
each city's 32 nearest neighbours ranked by a small MLP.

Everything runs on the device through ``libtspmi.so``: the 12-feature builder, the
12 -> 128 -> 128 -> 1 MLP (bf16 weights, fp32 accumulation, MFMA) forward and backward, Adam on
fp32 master weights, and the top-k selection that feeds the packed candidate 2-opt kernel.
Training labels are the edges of locally optimal tours produced by the batched memetic GA itself
(``BatchedMemeticGA``: tournament, PMX, inversion, survival and candidate 2-opt to the local
optimum).  Plain k-NN lists and the full neighbourhood stay available as the fallback and as the
judge of the tour gap the learned lists cost.

The reference has no learned component; its candidate lists are k-nearest neighbours
(src/distances/matrix.py:140-211, used by two_opt.py:117-149).
"""

from __future__ import annotations

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
            weight_decay: float = 0.0, log_every: int = 50) -> list[float]:
    """Minibatch Adam on (X, y); minibatch rows drawn on the host with a seeded generator."""
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
      _native.adam(self.theta, self.m, self.v, grad, self.step, lr, 0.9, 0.999, 1e-8, weight_decay, self.wbf)
      if log_every and (it % log_every == 0 or it == steps - 1):
        losses.append(float(grad[NPARAM].item()))
    self.history.extend(losses)
    return losses

  def state(self) -> np.ndarray:
    return self.theta.cpu().numpy()

  def save(self, path: str) -> None:
    np.savez(path, theta=self.state(), step=self.step)

  @classmethod
  def load(cls, path: str) -> "LearnedCandidateScorer":
    data = np.load(path, allow_pickle=False)
    s = cls(theta=data["theta"])
    s.step = int(data["step"])
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
