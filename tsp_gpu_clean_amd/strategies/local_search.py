"""Local search strategies: the HIP 2-opt drop-in for ``LocalSearchGPUTourTwoOpt``
(reference src/algorithms/strategies/local_search/two_opt.py:631-899) and ``LocalSearchNoOp``."""

from __future__ import annotations

from typing import Any

from .. import _native
from ..route_helpers import ensure_routes_batch, is_gpu_array, restore_routes_shape, to_device_int32
from ..context import coordinate_candidates_pay
from ..timing import time_strategy
from .base import StrategyBase

TSPMI_MODE_BEST = 0
TSPMI_MODE_FIRST = 1
TSPMI_MODE_MULTI = 2

#: Accepted modes.  The reference GPU kernel families "batch" and "naive" both apply the single
#: steepest move per pass; here they share the steepest kernel, whose ties follow the reference
#: CPU scan order (lowest (i, j)).  "best"/"first" name the CPU acceptance rules directly.
#: "multi_move" applies every mutually disjoint improving move of a pass by the exact CPU rule
#: (two_opt.py:491-563: steepest first, ties in scan order); the reference's GPU kernel of that
#: name offers only one candidate per thread (two_opt.py:649-655), so it is the CPU rule that
#: this device kernel reproduces bit for bit.
_MODES: dict[str, int] = {"batch": TSPMI_MODE_BEST, "naive": TSPMI_MODE_BEST, "best": TSPMI_MODE_BEST,
                          "first": TSPMI_MODE_FIRST, "multi_move": TSPMI_MODE_MULTI}


class LocalSearchGPUTourTwoOpt(StrategyBase):
  """Population-batched 2-opt on the MI355X; one launch refines every tour for all its passes."""

  def __init__(
    self,
    max_iterations: int = 10,
    mode: str = "batch",
    threads_per_block: int = 256,
    max_matrix_dimension: int | None = None,
    max_shared_tour_cities: int = 13_000,
    candidate_k: int | None = None,
    seed: Any = None,
    candidate_scorer: Any = None,
  ) -> None:
    """``candidate_scorer`` (a ``LearnedCandidateScorer``) replaces the k-nearest lists with the
    k best-scored of each city's 32 nearest neighbours; None keeps the reference's k-NN lists."""
    super().__init__(seed)
    if mode not in _MODES:
      raise ValueError(f"Invalid 2-opt GPU mode {mode!r}; expected one of {sorted(_MODES)}")
    if max_iterations < 0:
      raise ValueError(f"max_iterations must be >= 0, got {max_iterations}")
    self.mode = mode
    self.max_iterations = max_iterations
    self.threads_per_block = threads_per_block
    self.max_matrix_dimension = max_matrix_dimension
    self.max_shared_tour_cities = max_shared_tour_cities
    self.candidate_k = candidate_k
    self.candidate_scorer = candidate_scorer
    if candidate_scorer is not None and not (candidate_k and 1 <= candidate_k <= 32):
      raise ValueError("a candidate scorer needs candidate_k in [1, 32]")
    self.last_stats: tuple[int, int, int] | None = None
    self._stats = None
    self._full_ws = None  # workspace of the renumbered full scan
    self._cand_ws = None  # workspace of the renumbered candidate scan

  def _warmup_impl(self, context) -> None:
    self.improve_batch(self._warmup_batch(context, rows=1), context)

  def _uses_candidates(self, context, n: int) -> bool:
    """Candidate lists only for cyclic TSP tours spanning the whole instance (two_opt.py:711-736)."""
    if self.candidate_k is None or self.candidate_k <= 0:
      return False
    if bool(getattr(context, "is_closed", False)) or not bool(getattr(context, "return_to_depot", True)):
      return False
    return n == int(context.dimension)

  def improve(self, route, context):
    return self.improve_batch(route, context)

  @time_strategy
  def improve_batch(self, routes, context):
    """Refine (n,) or (N, n) tours; device in -> device out without a host sync."""
    torch = _native.require_device()
    batch, was_1d = ensure_routes_batch(routes)
    num, n = int(batch.shape[0]), int(batch.shape[1])
    on_gpu = is_gpu_array(batch)
    if n < 4 or self.max_iterations <= 0 or num == 0:
      out = batch.clone() if on_gpu else batch.copy()
      return restore_routes_shape(out, was_1d)
    if not bool(getattr(context, "return_to_depot", True)):
      raise NotImplementedError("open-path (return_to_depot=False) 2-opt is not implemented on the device")
    tours, _, up = to_device_int32(batch, copy=True)
    self.h2d_bytes += up
    dist = context.get_gpu_distances()
    cand = table = None
    if self._uses_candidates(context, n):
      if self.candidate_scorer is not None:
        cand, table = context.get_gpu_learned_candidates(self.candidate_scorer, int(self.candidate_k))
      else:
        cand = context.get_gpu_candidate_list(int(self.candidate_k))
    if self._stats is None:
      self._stats = torch.zeros(3, dtype=torch.int64, device="cuda")
    else:
      self._stats.zero_()
    large = n > (_native.CAND_LDS_MAX_N if cand is not None else _native.FULL_LDS_MAX_N)
    if large and _MODES[self.mode] != TSPMI_MODE_BEST:
      raise ValueError(f"2-opt mode {self.mode!r} at n={n}: above the shared-memory limits only the steepest "
                       "('batch', 'naive', 'best') scan is implemented")
    if cand is not None and _MODES[self.mode] == TSPMI_MODE_BEST and 1 <= cand.shape[1] <= 32:
      if table is None:
        table = context.get_gpu_candidate_table(int(self.candidate_k))
      ixy = context.get_gpu_int_coordinates()  # integer instances: the integer distance form
      pay = coordinate_candidates_pay(n, int(tours.shape[0]), integer=ixy is not None)
      xy = context.get_gpu_euc2d_coordinates() if pay else None
      if large:  # the tour arrays in a global workspace (two_opt.py:75-94)
        _native.two_opt_large(tours, dist, xy, table, int(cand.shape[1]), self.max_iterations, self._stats)
      elif xy is not None:  # same tours; distances from the coordinates
        ws = None
        if ixy is not None:  # large tours: the scan on the instance renumbered in Hilbert order
          need = _native.packed_euc2d_ws_bytes(n, int(cand.shape[1]))
          if self._cand_ws is None or self._cand_ws.numel() < need:
            self._cand_ws = torch.empty(need, dtype=torch.uint8, device="cuda")
          ws = self._cand_ws
        _native.two_opt_packed_euc2d(tours, ixy if ixy is not None else xy, table, int(cand.shape[1]),
                                     self.max_iterations, self._stats, ws=ws)
      else:
        _native.two_opt_packed(tours, dist, table, int(cand.shape[1]), self.max_iterations, self._stats)
    elif large and cand is None:
      _native.two_opt_large(tours, dist, context.get_gpu_euc2d_coordinates(), None, 0, self.max_iterations, self._stats)
    else:
      xy = None
      if cand is None and _MODES[self.mode] == TSPMI_MODE_BEST and n <= _native.FULL_EUC2D_MAX_N:
        xy = context.get_gpu_euc2d_coordinates()
      if xy is not None:  # same tours; 13 rows in 16 take their distances from the coordinates
        # the scan on the instance renumbered in Hilbert order (same moves), in a workspace kept per n
        need = _native.full_euc2d_ws_bytes(n)
        if self._full_ws is None or self._full_ws.numel() < need:
          self._full_ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        _native.two_opt_full_euc2d(tours, dist, xy, self.max_iterations, self._stats, ws=self._full_ws)
      else:
        _native.two_opt(tours, dist, cand, _MODES[self.mode], self.max_iterations, self._stats)
    self.kernel_launches += 1
    if on_gpu:
      return restore_routes_shape(tours, was_1d)
    host = tours.cpu().numpy()
    self.d2h_bytes += int(host.nbytes)
    return restore_routes_shape(host, was_1d)

  def read_stats(self) -> tuple[int, int, int]:
    """(deltas evaluated, moves applied, passes) of the last call (synchronises)."""
    if self._stats is None:
      return (0, 0, 0)
    s = self._stats.cpu().tolist()
    self.last_stats = (int(s[0]), int(s[1]), int(s[2]))
    return self.last_stats


class LocalSearchNoOp(StrategyBase):
  """Identity local search (reference local_search/__init__.py LocalSearchNoOp)."""

  def improve(self, route, context):
    return route.clone() if is_gpu_array(route) else route.copy()

  def improve_batch(self, routes, context):
    return routes.clone() if is_gpu_array(routes) else routes.copy()
