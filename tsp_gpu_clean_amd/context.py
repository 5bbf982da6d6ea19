"""ProblemContext: one run's generator plus the device-resident distance and candidate tables.

Mirrors the reference's ``ProblemContext`` (src/protocols/problem_context.py:75-616): a PCG64
``rng`` that every strategy draws from, lazily built and cached distance matrix and k-NN tables,
and an ``h2d_bytes`` counter of physical uploads.  Here the tables are built on the device by
``tspmi_dist_euc2d`` / ``tspmi_knn``; host copies are downloads of those device tables.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from . import _native
from .errors import VRAMInsufficientError
from .problem import Problem
from .rng import make_rng

#: Largest |D| for which every int32 2-opt delta is exact (a + b - c - d with 0 <= D < 2^29).
MAX_DISTANCE = (1 << 29) - 1
DEFAULT_MAX_MATRIX_DIMENSION = 10_000


def coordinate_candidates_pay(n: int, tours: int | None = None, integer: bool = False) -> bool:
  """Whether the candidate-list 2-opt should compute EUC_2D distances from coordinates instead of
  gathering them from the matrix.  Measured on an MI355X with identical tours (k = 20):
  * integer coordinates (`integer`, the int32 form with the coordinates and node records in LDS):
    20 passes at P = 1024 take 0.16 / 0.19 / 0.28 / 0.47 ms against the matrix's 0.18 / 0.27 /
    0.44 / 1.02 at n = 200 / 300 / 500 / 1000; at P = 256, 0.094 / 0.125 / 0.187 against 0.081 /
    0.168 / 0.349 at n = 200 / 500 / 1000; at P = 16 (10 passes) 0.094 / 0.151 against 0.144 /
    0.358 at n = 1000 / 2000 -- hence n > 256;
  * FP64 coordinates, many tours (more than CUs, the throughput regime; P = 1024): the matrix form
    wins at n = 500 (2.4e11 vs 1.9e11 deltas/s) and loses from n = 1000 on (1.06 vs 0.90 ms per
    20-pass launch at n = 1000, 1.3x at P = 4096; 2.5x at n = 2000, where the n x n int32 matrix
    has outgrown an XCD's 4 MB L2) -- hence n > 768;
  * FP64 coordinates, a few tours (one per CU or fewer, latency-bound): the forms run within 10 %
    of each other until the matrix leaves the caches altogether (n = 10000, k = 8, 256 tours, 200
    passes: 9.4 vs 35.2 ms) -- hence n > 4096."""
  if integer:
    return n > 256
  if tours is not None and tours > _cu_count():
    return n > 768
  return n > 4096


#: Largest per-axis coordinate span for which the integer EUC_2D distance form equals the FP64 one.
INT_COORD_SPAN = 1 << 20


def int_coordinates(xy) -> np.ndarray | None:
  """(n, 2) int32 copy of EUC_2D coordinates for the kernels' integer distance form, or None when
  it does not apply: every coordinate must be an integer (|v| <= 2^30) and each axis must span at
  most 2^20, where that form gives the FP64 nint exactly."""
  a = np.asarray(xy, dtype=np.float64)
  if a.ndim != 2 or a.shape[1] != 2 or a.shape[0] == 0 or not np.isfinite(a).all():
    return None
  if not (a == np.rint(a)).all() or np.abs(a).max() > 2.0**30:
    return None
  if (a.max(axis=0) - a.min(axis=0)).max() > INT_COORD_SPAN:
    return None
  return np.ascontiguousarray(a.astype(np.int32))


def _cu_count() -> int:
  try:
    import torch

    return int(torch.cuda.get_device_properties(0).multi_processor_count)
  except Exception:  # no device visible: the few-tour rule applies
    return 1 << 30


def explicit_int32_distances(dist) -> np.ndarray:
  """An explicit matrix as int32, checked before the cast: integral values in [0, MAX_DISTANCE],
  so that neither a wrap of a large integer nor the truncation of a fraction can slip under the
  guard that keeps every int32 2-opt delta exact."""
  a = np.asarray(dist)
  if a.size:
    if a.dtype.kind == "b" or a.dtype.kind not in "iuf":
      raise ValueError(f"explicit distances must be numeric, got dtype {a.dtype}")
    if a.dtype.kind == "f" and not np.all(np.isfinite(a)):
      raise ValueError("explicit distances must be finite")
    lo, hi = a.min(), a.max()
    if lo < 0:
      raise ValueError("explicit distances must be non-negative")
    if hi > MAX_DISTANCE:
      raise ValueError(f"explicit distance {hi} exceeds {MAX_DISTANCE}: int32 2-opt deltas would not be exact")
    if a.dtype.kind == "f" and not np.array_equal(a, np.floor(a)):
      raise ValueError("explicit distances must be integers")
  return np.ascontiguousarray(a, dtype=np.int32)


def reference_knn_rows(rows: np.ndarray, row_ids: np.ndarray, k: int) -> np.ndarray:
  """The reference's k-NN rule on distance rows `rows` ((m, n), row r is city row_ids[r]): the
  city itself masked with the int64 maximum, ``np.argpartition(block, k, axis=1)[:, :k]``, then
  ``np.lexsort((cols, dist))`` (src/distances/matrix.py:189, 204-209).  On a row whose k-th and
  (k+1)-th distances tie, argpartition decides which tied city enters the list; the device table
  keeps (distance, index) order, which equals this rule on every other row, so only the flagged
  rows are rebuilt here, with the same NumPy call the reference makes on the same host."""
  block = np.asarray(rows, dtype=np.int64).copy()
  m = block.shape[0]
  local = np.arange(m)
  block[local, np.asarray(row_ids, dtype=np.int64)] = np.iinfo(np.int64).max
  cols = np.argpartition(block, k, axis=1)[:, :k]
  cdist = block[local[:, None], cols]
  order = np.lexsort((cols, cdist), axis=1)
  return np.ascontiguousarray(cols[local[:, None], order], dtype=np.int32)


class ProblemContext:
  def __init__(
    self,
    problem: Problem,
    xp: Any = None,
    seed: Any = None,
    max_matrix_dimension: int = DEFAULT_MAX_MATRIX_DIMENSION,
    candidate_cache: dict | None = None,
  ) -> None:
    self.xp = xp
    self.problem = problem
    self.dimension = int(problem.dimension)
    self.capacity = problem.capacity
    self.max_matrix_dimension = max_matrix_dimension
    self.rng: np.random.Generator = make_rng(seed)
    self.seed = seed
    self.h2d_bytes = 0
    self.d2h_bytes = 0
    self._gpu_distances = None
    self._cpu_distances: np.ndarray | None = None
    self._gpu_coordinates = None
    self._gpu_int_coordinates = None  # False once known not to apply
    self._gpu_candidates: dict[int, Any] = {}
    self._cpu_candidates: dict[int, np.ndarray] = candidate_cache if candidate_cache is not None else {}
    self._candidate_ties: dict[int, int] = {}
    self.max_distance: int | None = None

  # -- topology (problem_context.py:219-237) ---------------------------------------------------
  @property
  def return_to_depot(self) -> bool:
    return bool(self.problem.constraints.return_to_depot)

  @property
  def is_closed(self) -> bool:
    return False  # TSP/ATSP permutations are open (no depot-bounded routes)

  # -- device tables ----------------------------------------------------------------------------
  def _check_vram(self, nbytes: int, what: str) -> None:
    torch = _native.require_device()
    free, _ = torch.cuda.mem_get_info()
    if nbytes > free:
      raise VRAMInsufficientError(f"{what} needs {nbytes} B but only {free} B of device memory are free")

  def get_gpu_coordinates(self):
    torch = _native.require_device()
    if self._gpu_coordinates is None and self.problem.coordinates is not None:
      xy = np.ascontiguousarray(self.problem.coordinates, dtype=np.float64)
      self._gpu_coordinates = torch.from_numpy(xy).to("cuda")
      self.h2d_bytes += int(xy.nbytes)
    return self._gpu_coordinates

  def get_gpu_euc2d_coordinates(self):
    """Device (n, 2) float64 coordinates when the distances are EUC_2D of them, else None."""
    if self.problem.edge_type != "EUC_2D" or self.problem.coordinates is None or self.problem.distances is not None:
      return None
    return self.get_gpu_coordinates()

  def get_gpu_int_coordinates(self):
    """Device (n, 2) int32 coordinates for the integer form of the EUC_2D distance (same values,
    about half the work per distance), or None unless the distances are EUC_2D of integer
    coordinates spanning at most 2^20 per axis (``int_coordinates``)."""
    torch = _native.require_device()
    if self._gpu_int_coordinates is None:
      ixy = None
      if self.get_gpu_euc2d_coordinates() is not None:
        ixy = int_coordinates(self.problem.coordinates)
      self._gpu_int_coordinates = False if ixy is None else torch.from_numpy(ixy).to("cuda")
      if ixy is not None:
        self.h2d_bytes += int(ixy.nbytes)
    return self._gpu_int_coordinates if self._gpu_int_coordinates is not False else None

  def get_gpu_distances(self):
    """Dense (n, n) int32 matrix on the device, built once (problem_context.py:337-397)."""
    torch = _native.require_device()
    if self._gpu_distances is None:
      n = self.dimension
      self._check_vram(4 * n * n, f"dense distance matrix (n={n})")
      if self.problem.distances is not None:
        host = explicit_int32_distances(self.problem.distances)
        d = torch.from_numpy(host).to("cuda")
        self.h2d_bytes += int(host.nbytes)
        mx = int(host.max()) if host.size else 0
      else:
        xy = self.get_gpu_coordinates()
        d = torch.empty((n, n), dtype=torch.int32, device="cuda")
        mx_t = torch.zeros(1, dtype=torch.int32, device="cuda")
        if self.problem.edge_type == "EUC_2D":
          _native.dist_euc2d(xy, d, mx_t)
        else:
          _native.dist_metric(xy, self.problem.edge_type, d, mx_t)
        mx = int(mx_t.item())
      if mx > MAX_DISTANCE:
        raise ValueError(f"max distance {mx} exceeds {MAX_DISTANCE}: int32 2-opt deltas would not be exact")
      self.max_distance = mx
      self._gpu_distances = d
    return self._gpu_distances

  def get_cpu_distances(self) -> np.ndarray:
    """Host copy of the device matrix (problem_context.py:266-301)."""
    if self._cpu_distances is None:
      self._cpu_distances = self.get_gpu_distances().cpu().numpy()
    return self._cpu_distances

  def get_gpu_candidate_list(self, k: int = 20):
    """(n, min(k, n-1)) int32 k-NN table in (distance, index) order, built on device (problem_context.py:456-480)."""
    torch = _native.require_device()
    table = self._gpu_candidates.get(k)
    if table is None:
      n = self.dimension
      k_eff = min(k, n - 1) if n > 1 else 0
      cached = self._cpu_candidates.get(k)
      if cached is not None:
        table = torch.from_numpy(np.ascontiguousarray(cached, dtype=np.int32)).to("cuda")
        self.h2d_bytes += int(cached.nbytes)
      else:
        d = self.get_gpu_distances()
        table = torch.empty((n, k_eff), dtype=torch.int32, device="cuda")
        ties = 0
        if k_eff > 0:
          flags = torch.empty(n, dtype=torch.int32, device="cuda")
          _native.knn_flagged(d, k_eff, table, flags)
          rows = torch.nonzero(flags).flatten()  # host sync: once per (context, k)
          ties = int(rows.numel())
          if ties:
            fixed = reference_knn_rows(d.index_select(0, rows)[:, :n].cpu().numpy(), rows.cpu().numpy(), k_eff)
            self.d2h_bytes += ties * n * 4
            table.index_copy_(0, rows, torch.from_numpy(fixed).to("cuda"))
            self.h2d_bytes += int(fixed.nbytes)
        self._candidate_ties[k] = ties
      self._gpu_candidates[k] = table
    return table

  def get_gpu_candidate_table(self, k: int = 20):
    """Packed (n, KP, 2) int32 {neighbour, distance} table for the fast candidate 2-opt (k <= 32)."""
    torch = _native.require_device()
    key = ("packed", k)
    table = self._gpu_candidates.get(key)
    if table is None:
      cand = self.get_gpu_candidate_list(k)
      n, k_eff = cand.shape
      table = torch.empty((n, _native.candidate_lanes(max(1, k_eff)), 2), dtype=torch.int32, device="cuda")
      if k_eff > 0:
        _native.candidate_table(cand, self.get_gpu_distances(), table)
      self._gpu_candidates[key] = table
    return table

  def get_gpu_learned_candidates(self, scorer, k: int = 8):
    """(list (n, k), packed table) from a learned candidate scorer, built once per (scorer, k).

    The scorer ranks each city's 32 nearest neighbours; an instance with n <= 32 keeps its k-NN
    lists, since every other city is already a neighbour."""
    torch = _native.require_device()
    key = ("learned", k, id(scorer))
    hit = self._gpu_candidates.get(key)
    if hit is None:
      n = self.dimension
      if n <= 32:
        cand = self.get_gpu_candidate_list(k)
        hit = (cand, self.get_gpu_candidate_table(k))
      else:
        cand = scorer.candidates(self, k)
        table = torch.empty((n, _native.candidate_lanes(k), 2), dtype=torch.int32, device="cuda")
        _native.candidate_table(cand, self.get_gpu_distances(), table)
        hit = (cand, table)
      self._gpu_candidates[key] = hit
    return hit

  def get_cpu_candidate_list(self, k: int = 20) -> np.ndarray:
    cached = self._cpu_candidates.get(k)
    if cached is None:
      cached = self.get_gpu_candidate_list(k).cpu().numpy()
      self._cpu_candidates[k] = cached
    return cached

  def candidate_boundary_ties(self, k: int = 20) -> int:
    """Rows whose k-th and (k+1)-th distances tie: the rows rebuilt on the host by the
    reference's argpartition rule (reference_knn_rows)."""
    self.get_gpu_candidate_list(k)
    return int(self._candidate_ties.get(k, 0))
