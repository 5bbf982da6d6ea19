"""ctypes binding of ``libtspmi.so`` — the only way the package reaches the device.

Every wrapper takes torch CUDA (ROCm/HIP) tensors, checks shapes and dtypes on the host,
passes raw device pointers plus the current HIP stream, and maps the library status to an
exception: ``ValueError`` for a rejected argument, ``RuntimeError`` for a HIP failure and
``GPUUnavailableError`` when the library or the device is missing.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import Any

import numpy as np

from .errors import GPUUnavailableError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtspmi.so")

_lib: Any = None
_lock = threading.Lock()

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double

# name -> argtypes (all return int status)
_SIGNATURES: dict[str, list[Any]] = {
  "tspmi_version": [],
  "tspmi_device_count": [_P],
  "tspmi_probe_gather": [_P, _I, _I, _I, _P, _P],
  "tspmi_dist_euc2d": [_P, _I, _P, _I, _P, _P],
  "tspmi_dist_euc2d_i32": [_P, _I, _P, _I, _P, _P],
  "tspmi_knn": [_P, _I, _I, _I, _P, _P, _P],
  "tspmi_knn_flagged": [_P, _I, _I, _I, _P, _P, _P],
  "tspmi_tour_cost": [_P, _I, _I, _P, _I, _I, _P, _P],
  "tspmi_two_opt": [_P, _I, _I, _P, _I, _P, _I, _I, _I, _P, _P],
  "tspmi_two_opt_full_euc2d": [_P, _I, _I, _P, _I, _P, _I, _P, _P],
  "tspmi_two_opt_full_euc2d_ordered": [_P, _I, _I, _P, _I, _P, _I, _P, _P, ctypes.c_size_t, _P],
  "tspmi_two_opt_full_euc2d_ws_bytes": [_I],
  "tspmi_candidate_table": [_P, _I, _I, _P, _I, _P, _P],
  "tspmi_two_opt_packed": [_P, _I, _I, _P, _I, _P, _I, _I, _P, _P],
  "tspmi_two_opt_packed_cost": [_P, _I, _I, _P, _I, _P, _I, _I, _P, _P, _P],
  "tspmi_two_opt_elite": [_P, _I, _I, _P, _I, _P, _I, _I, _P, _P, _P, _P, _P],
  "tspmi_two_opt_elite_noalloc": [_P, _I, _I, _P, _I, _P, _I, _I, _P, _P, _P, _P, _P],
  "tspmi_two_opt_elite_xy": [_P, _I, _I, _P, _P, _I, _I, _P, _P, _P, _P, _P],
  "tspmi_tournament": [_P, _I, _I, _P, _P, _I, _I, _P, _P, _P],
  "tspmi_pmx": [_P, _I, _I, _P, _P, _P, _P],
  "tspmi_inversion": [_P, _I, _I, _P, _P, _P, _D, _P],
  "tspmi_generation": [_P, _P, _I, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P, _D, _P, _P, _P],
  "tspmi_draw_generations": [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P],
  "tspmi_rank_select": [_P, _I, _I, _P, _P],
  "tspmi_survive": [_P, _I, _I, _P, _I, _P, _I, _P, _P, _P, _P],
  "tspmi_gather_rows": [_P, _P, _I, _P, _P, _I, _P, _I, _P, _P, _P],
  "tspmi_scatter_rows": [_P, _P, _P, _I, _I, _P, _P, _P],
  "tspmi_random_init": [_P, _I, _P, _I, _P, _P],
  "tspmi_nn_init": [_P, _I, _P, _I, _P, _I, _P, _P, _P],
  "tspmi_batch_dist_euc2d": [_P, _P, _P, _I, _P, _P, _P],
  "tspmi_batch_knn": [_P, _P, _P, _I, _I, _P, _P, _P, _P],
  "tspmi_batch_random_init": [_P, _P, _P, _I, _P, _P, _P],
  "tspmi_batch_cost": [_P, _P, _P, _I, _P, _P, _P, _P, _P, _P],
  "tspmi_batch_generation": [_P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _D, _P, _I, _I, _P],
  "tspmi_batch_survival": [_P, _I, _P, _P, _P],
  "tspmi_batch_gather": [_P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P],
  "tspmi_batch_two_opt_packed": [_P, _P, _P, _I, _I, _P, _P, _P, _I, _I, _I, _I, _P, _P],
  "tspmi_batch_two_opt_packed_mixed": [_P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _P],
  "tspmi_batch_best": [_P, _I, _P, _I, _P, _P],
  "tspmi_batch_fused_generation": [_P, _I, _P, _I, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P,
                                   _P, _P, _P, _P, _D, _I, _I, _I, _I, _I, _I, _P, _P, _I, _P],
  "tspmi_nn_greedy_init": [_P, _I, _I, _P, _I, _P, _P],
  "tspmi_ox": [_P, _I, _I, _P, _P, _P, _P],
  "tspmi_dist_metric": [_P, _I, _I, _P, _I, _P, _P],
  "tspmi_swap": [_P, _I, _I, _P, _P, _P, _D, _P],
  "tspmi_two_opt_euc2d": [_P, _I, _I, _P, _I, _P, _P],
  "tspmi_two_opt_packed_euc2d": [_P, _I, _I, _P, _P, _I, _I, _P, _P],
  "tspmi_two_opt_packed_euc2d_cost": [_P, _I, _I, _P, _P, _I, _I, _P, _P, _P],
  "tspmi_two_opt_packed_euc2d_log": [_P, _I, _I, _P, _P, _I, _I, _P, _P, _P, _I, _P],
  "tspmi_two_opt_packed_euc2d_i32": [_P, _I, _I, _P, _P, _I, _I, _P, _P, _P, _I, _P],
  "tspmi_two_opt_packed_euc2d_i32_ordered": [_P, _I, _I, _P, _P, _I, _I, _P, _P, _P, _I, _P, ctypes.c_size_t, _P],
  "tspmi_two_opt_packed_euc2d_ws_bytes": [_I, _I],
  "tspmi_wall_clock_khz": [_P],
  "tspmi_two_opt_large": [_P, _I, _I, _P, _I, _P, _P, _I, _I, _P, _P, _P, _P],
  "tspmi_scorer_param_count": [],
  "tspmi_scorer_features": [_P, _I, _P, _P, _I, _P, _P],
  "tspmi_scorer_labels": [_P, _I, _P, _P, _P, _P],
  "tspmi_mlp_forward": [_P, _I, _P, _P, _P],
  "tspmi_mlp_grad": [_P, _P, _I, _P, _P, _P, _P, _P],
  "tspmi_mlp_work_sizes": [_I, _P, _P],
  "tspmi_adam": [_P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                 _P, _P],
  "tspmi_to_bf16": [_P, _I, _P, _P],
  "tspmi_scorer_topk": [_P, _P, _I, _I, _P, _P],
}

EXPORTED = tuple(_SIGNATURES) + ("tspmi_last_error",)

TSPMI_EINVAL = -1
TSPMI_EHIP = -2
TSPMI_ENODEV = -3


def load_library(path: str | None = None) -> Any:
  """Load and type the shared library (no device work)."""
  global _lib
  with _lock:
    if _lib is not None and path is None:
      return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
      raise GPUUnavailableError(
        f"libtspmi.so not found at {p}; build it with `python -c 'import __graft_entry__ as g; g.build()'`"
      )
    try:
      lib = ctypes.CDLL(p)
    except OSError as exc:  # missing HIP runtime etc.
      raise GPUUnavailableError(f"cannot load {p}: {exc}") from exc
    for name, argtypes in _SIGNATURES.items():
      fn = getattr(lib, name)
      fn.argtypes = argtypes
      fn.restype = ctypes.c_int
    lib.tspmi_last_error.argtypes = []
    lib.tspmi_last_error.restype = ctypes.c_char_p
    lib.tspmi_two_opt_full_euc2d_ws_bytes.restype = ctypes.c_size_t
    lib.tspmi_two_opt_packed_euc2d_ws_bytes.restype = ctypes.c_size_t
    if path is None:
      _lib = lib
    return lib


def _check(rc: int, fn: str) -> None:
  if rc == 0:
    return
  msg = (_lib.tspmi_last_error() or b"").decode(errors="replace")
  if rc == TSPMI_EINVAL:
    raise ValueError(f"{fn}: {msg}")
  if rc == TSPMI_ENODEV:
    raise GPUUnavailableError(f"{fn}: {msg}")
  raise RuntimeError(f"{fn} failed ({rc}): {msg}")


def require_device():
  """Return torch with a usable HIP device, else raise GPUUnavailableError (never falls back)."""
  try:
    import torch
  except Exception as exc:  # pragma: no cover
    raise GPUUnavailableError(f"torch unavailable: {exc}") from exc
  if not torch.cuda.is_available():
    raise GPUUnavailableError("no HIP device visible to torch; the GPU strategies never fall back to CPU")
  load_library()
  return torch


def _stream(torch_mod) -> int:
  return int(torch_mod.cuda.current_stream().cuda_stream)


def _ptr(t) -> int | None:
  if t is None:
    return None
  return int(t.data_ptr())


def _need(t, dtype, name: str, ndim: int | None = None) -> None:
  import torch

  if not isinstance(t, torch.Tensor) or not t.is_cuda:
    raise ValueError(f"{name} must be a device tensor")
  if t.dtype != dtype:
    raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
  if not t.is_contiguous():
    raise ValueError(f"{name} must be contiguous")
  if ndim is not None and t.dim() != ndim:
    raise ValueError(f"{name} must be {ndim}-D, got shape {tuple(t.shape)}")


# ---------------------------------------------------------------------------------------------
# Thin typed wrappers (device tensors in, enqueue on the current stream, no host sync)
# ---------------------------------------------------------------------------------------------


def dist_euc2d(xy, D, max_out=None) -> None:
  torch = require_device()
  _need(xy, torch.float64, "xy", 2)
  _need(D, torch.int32, "D", 2)
  n = xy.shape[0]
  if xy.shape[1] != 2 or D.shape[0] < n or D.shape[1] < n:
    raise ValueError(f"dist_euc2d: xy {tuple(xy.shape)} vs D {tuple(D.shape)}")
  if max_out is not None:
    _need(max_out, torch.int32, "max_out")
  _check(_lib.tspmi_dist_euc2d(_ptr(xy), n, _ptr(D), D.shape[1], _ptr(max_out), _stream(torch)), "tspmi_dist_euc2d")


def dist_euc2d_i32(xy, D, bad_out=None) -> None:
  """EUC_2D matrix from int32 coordinates by the integer distance form; pairs more than 2^21 apart
  on an axis get -1 and are counted in ``bad_out`` (int32 (1,))."""
  torch = require_device()
  _need(xy, torch.int32, "xy", 2)
  _need(D, torch.int32, "D", 2)
  n = xy.shape[0]
  if xy.shape[1] != 2 or D.shape[0] < n or D.shape[1] < n:
    raise ValueError(f"dist_euc2d_i32: xy {tuple(xy.shape)} vs D {tuple(D.shape)}")
  if bad_out is not None:
    _need(bad_out, torch.int32, "bad_out")
  _check(_lib.tspmi_dist_euc2d_i32(_ptr(xy), n, _ptr(D), D.shape[1], _ptr(bad_out), _stream(torch)),
         "tspmi_dist_euc2d_i32")


def probe_gather(rows, blocks: int, iters: int, out) -> None:
  """Random 4-byte gathers inside the 4 KB rows of `rows` ((nrows, 1024) int32); calibration only."""
  torch = require_device()
  _need(rows, torch.int32, "rows", 2)
  _need(out, torch.int32, "out")
  if rows.shape[1] != 1024:
    raise ValueError(f"probe_gather: rows must be (nrows, 1024), got {tuple(rows.shape)}")
  _check(_lib.tspmi_probe_gather(_ptr(rows), rows.shape[0], blocks, iters, _ptr(out), _stream(torch)),
         "tspmi_probe_gather")


METRICS = {"EUC_2D": 0, "CEIL_2D": 1, "MAN_2D": 2, "MAX_2D": 3, "ATT": 4, "GEO": 5}


def dist_metric(xy, metric: str, D, max_out=None) -> None:
  """Dense int32 matrix of a TSPLIB 2D metric (EUC_2D, CEIL_2D, MAN_2D, MAX_2D, ATT, GEO)."""
  torch = require_device()
  _need(xy, torch.float64, "xy", 2)
  _need(D, torch.int32, "D", 2)
  if metric not in METRICS:
    raise ValueError(f"unsupported metric {metric!r}; expected one of {sorted(METRICS)}")
  n = xy.shape[0]
  if xy.shape[1] != 2 or D.shape[0] < n or D.shape[1] < n:
    raise ValueError(f"dist_metric: xy {tuple(xy.shape)} vs D {tuple(D.shape)}")
  if max_out is not None:
    _need(max_out, torch.int32, "max_out")
  _check(_lib.tspmi_dist_metric(_ptr(xy), n, METRICS[metric], _ptr(D), D.shape[1], _ptr(max_out), _stream(torch)),
         "tspmi_dist_metric")


def knn(D, k: int, out, ties_out=None) -> None:
  torch = require_device()
  _need(D, torch.int32, "D", 2)
  _need(out, torch.int32, "out", 2)
  n = D.shape[0]
  if out.shape != (n, k):
    raise ValueError(f"knn: out shape {tuple(out.shape)} != ({n}, {k})")
  if ties_out is not None:
    _need(ties_out, torch.int32, "ties_out")
  _check(_lib.tspmi_knn(_ptr(D), n, D.shape[1], k, _ptr(out), _ptr(ties_out), _stream(torch)), "tspmi_knn")


def knn_flagged(D, k: int, out, tie_flags) -> None:
  """k-NN table in (distance, index) order plus a per-row flag of k-th/(k+1)-th distance ties."""
  torch = require_device()
  _need(D, torch.int32, "D", 2)
  _need(out, torch.int32, "out", 2)
  _need(tie_flags, torch.int32, "tie_flags", 1)
  n = D.shape[0]
  if out.shape != (n, k) or tie_flags.shape[0] != n:
    raise ValueError(f"knn_flagged: out {tuple(out.shape)} / flags {tuple(tie_flags.shape)} for n={n}, k={k}")
  _check(_lib.tspmi_knn_flagged(_ptr(D), n, D.shape[1], k, _ptr(out), _ptr(tie_flags), _stream(torch)),
         "tspmi_knn_flagged")


def tour_cost(tours, D, out, closed: bool = True) -> None:
  torch = require_device()
  _need(tours, torch.int32, "tours", 2)
  _need(D, torch.int32, "D", 2)
  _need(out, torch.float64, "out", 1)
  P, n = tours.shape
  if out.shape[0] != P:
    raise ValueError("tour_cost: out length != P")
  _check(_lib.tspmi_tour_cost(_ptr(tours), P, n, _ptr(D), D.shape[1], int(bool(closed)), _ptr(out), _stream(torch)),
         "tspmi_tour_cost")


def two_opt(tours, D, cand=None, mode: int = 0, max_iter: int = 10, stats=None) -> None:
  """2-opt passes in place; max_iter < 0 runs to the local optimum (symmetric D only: on an
  asymmetric matrix the moves can cycle, so pass a bound)."""
  torch = require_device()
  _need(tours, torch.int32, "tours", 2)
  _need(D, torch.int32, "D", 2)
  k = 0
  if cand is not None:
    _need(cand, torch.int32, "cand", 2)
    k = cand.shape[1]
  if stats is not None:
    _need(stats, torch.int64, "stats")
    if stats.numel() < 3:
      raise ValueError("two_opt: stats needs 3 int64 slots")
  P, n = tours.shape
  _check(_lib.tspmi_two_opt(_ptr(tours), P, n, _ptr(D), D.shape[1], _ptr(cand), k, int(mode), int(max_iter),
                            _ptr(stats), _stream(torch)), "tspmi_two_opt")


def candidate_lanes(k: int) -> int:
  """Row width KP of the packed candidate table for k candidates."""
  return 8 if k <= 8 else (16 if k <= 16 else 32)


def candidate_table(cand, D, out) -> None:
  torch = require_device()
  _need(cand, torch.int32, "cand", 2)
  _need(D, torch.int32, "D", 2)
  _need(out, torch.int32, "out", 3)
  n, k = cand.shape
  if out.shape != (n, candidate_lanes(k), 2):
    raise ValueError(f"candidate_table: out shape {tuple(out.shape)} != ({n}, {candidate_lanes(k)}, 2)")
  _check(_lib.tspmi_candidate_table(_ptr(cand), n, k, _ptr(D), D.shape[1], _ptr(out), _stream(torch)),
         "tspmi_candidate_table")


def two_opt_packed(tours, D, table, k: int, max_iter: int = 10, stats=None, cost=None) -> None:
  """Candidate-list best-improvement 2-opt on a packed table; ``cost`` (float64 (P,), optional)
  receives each final tour's closed cost from the same launch."""
  torch = require_device()
  _need(tours, torch.int32, "tours", 2)
  _need(D, torch.int32, "D", 2)
  _need(table, torch.int32, "table", 3)
  if stats is not None:
    _need(stats, torch.int64, "stats")
  P, n = tours.shape
  if table.shape[0] != n or table.shape[1] != candidate_lanes(k):
    raise ValueError("two_opt_packed: table does not match n / k")
  if cost is not None:
    _need(cost, torch.float64, "cost", 1)
    if cost.shape[0] < P:
      raise ValueError("two_opt_packed: cost shorter than the number of tours")
  _check(_lib.tspmi_two_opt_packed_cost(_ptr(tours), P, n, _ptr(D), D.shape[1], _ptr(table), int(k), int(max_iter),
                                        _ptr(stats), _ptr(cost), _stream(torch)), "tspmi_two_opt_packed_cost")


ELITE_SYNC_WORDS = 528  # int64 words of the elite 2-opt workspace per tour
ELITE_ERROR_WORD = 512  # its barrier-timeout word


def two_opt_elite(tours, D, table, k: int, max_iter: int, stats, cost, sync, best=None, ixy=None) -> None:
  """Elite-set candidate 2-opt with final costs: :func:`two_opt_packed` semantics, one cluster of
  workgroups per tour when the tours leave CUs idle.  ``sync`` is an int64 device workspace of
  ``ELITE_SYNC_WORDS * P`` words, zeroed before first use; :func:`elite_failed` reports a cluster
  whose barrier timed out (its tour and cost are then left unchanged).  ``best`` (int32 (1,),
  optional) receives the index of the first minimum cost.  With ``ixy`` (int32 (n, 2) EUC_2D
  coordinates within the integer form's span) the distances come from the coordinates
  (tspmi_two_opt_elite_xy; same results) and ``D`` is not read."""
  torch = require_device()
  _need(tours, torch.int32, "tours", 2)
  if ixy is None:
    _need(D, torch.int32, "D", 2)
  else:
    _need(ixy, torch.int32, "ixy", 2)
  _need(table, torch.int32, "table", 3)
  _need(sync, torch.int64, "sync", 1)
  if stats is not None:
    _need(stats, torch.int64, "stats")
  P, n = tours.shape
  if table.shape[0] != n or table.shape[1] != candidate_lanes(k):
    raise ValueError("two_opt_elite: table does not match n / k")
  if cost is not None:
    _need(cost, torch.float64, "cost", 1)
    if cost.shape[0] < P:
      raise ValueError("two_opt_elite: cost shorter than the number of tours")
  if sync.shape[0] < ELITE_SYNC_WORDS * P:
    raise ValueError(f"two_opt_elite: sync needs {ELITE_SYNC_WORDS * P} words")
  if best is not None:
    _need(best, torch.int32, "best", 1)
    if cost is None:
      raise ValueError("two_opt_elite: best needs cost")
  if ixy is not None:
    if ixy.shape != (n, 2):
      raise ValueError("two_opt_elite: ixy must be (n, 2)")
    _check(_lib.tspmi_two_opt_elite_xy(_ptr(tours), P, n, _ptr(ixy), _ptr(table), int(k), int(max_iter),
                                       _ptr(stats), _ptr(cost), _ptr(best), _ptr(sync), _stream(torch)),
           "tspmi_two_opt_elite_xy")
    return
  _check(_lib.tspmi_two_opt_elite(_ptr(tours), P, n, _ptr(D), D.shape[1], _ptr(table), int(k), int(max_iter),
                                  _ptr(stats), _ptr(cost), _ptr(best), _ptr(sync), _stream(torch)),
         "tspmi_two_opt_elite")


ELITE_STATIC_TOURS = 128  # TSPMI_ELITE_STATIC_TOURS: capacity of the library's static elite workspace


def two_opt_elite_noalloc(tours, D, table, k: int, max_iter: int, stats, cost, best=None, failed=None) -> None:
  """:func:`two_opt_elite` on the library's static workspace (no allocation, P <= 128); ``failed``
  (int32 (1,), optional) receives the number of tours whose cluster barrier timed out."""
  torch = require_device()
  _need(tours, torch.int32, "tours", 2)
  _need(D, torch.int32, "D", 2)
  _need(table, torch.int32, "table", 3)
  P, n = tours.shape
  if table.shape[0] != n or table.shape[1] != candidate_lanes(k):
    raise ValueError("two_opt_elite_noalloc: table does not match n / k")
  for t, name in ((stats, "stats"), (cost, "cost"), (best, "best"), (failed, "failed")):
    if t is not None:
      _need(t, torch.float64 if name == "cost" else (torch.int64 if name == "stats" else torch.int32), name)
  if cost is not None and cost.shape[0] < P:
    raise ValueError("two_opt_elite_noalloc: cost shorter than the number of tours")
  _check(_lib.tspmi_two_opt_elite_noalloc(_ptr(tours), P, n, _ptr(D), D.shape[1], _ptr(table), int(k), int(max_iter),
                                          _ptr(stats), _ptr(cost), _ptr(best), _ptr(failed), _stream(torch)),
         "tspmi_two_opt_elite_noalloc")


def elite_failed(sync) -> bool:
  """True when a cluster barrier of :func:`two_opt_elite` timed out (synchronises the device)."""
  return bool((sync.view(-1, ELITE_SYNC_WORDS)[:, ELITE_ERROR_WORD] != 0).any().item())


def tournament(pop, fitness, cand, out, winners_out=None) -> None:
  torch = require_device()
  _need(pop, torch.int32, "pop", 2)
  _need(fitness, torch.float64, "fitness", 1)
  _need(cand, torch.int32, "cand", 2)
  _need(out, torch.int32, "out", 2)
  P, n = pop.shape
  rows, k = cand.shape
  if out.shape != (rows, n) or fitness.shape[0] != P:
    raise ValueError("tournament: shape mismatch")
  _check(_lib.tspmi_tournament(_ptr(pop), P, n, _ptr(fitness), _ptr(cand), k, rows, _ptr(out), _ptr(winners_out),
                               _stream(torch)), "tspmi_tournament")


def pmx(parents, cut_lo, cut_hi, out) -> None:
  torch = require_device()
  _need(parents, torch.int32, "parents", 2)
  _need(out, torch.int32, "out", 2)
  P, n = parents.shape
  if P >= 2:
    _need(cut_lo, torch.int32, "cut_lo", 1)
    _need(cut_hi, torch.int32, "cut_hi", 1)
    if cut_lo.shape[0] < P // 2 or cut_hi.shape[0] < P // 2:
      raise ValueError("pmx: need P // 2 cut pairs")
  _check(_lib.tspmi_pmx(_ptr(parents), P, n, _ptr(cut_lo), _ptr(cut_hi), _ptr(out), _stream(torch)), "tspmi_pmx")


def inversion(routes, probs, pos1, pos2, rate: float) -> None:
  torch = require_device()
  _need(routes, torch.int32, "routes", 2)
  _need(probs, torch.float64, "probs", 1)
  _need(pos1, torch.int32, "pos1", 1)
  _need(pos2, torch.int32, "pos2", 1)
  P, n = routes.shape
  if min(probs.shape[0], pos1.shape[0], pos2.shape[0]) < P:
    raise ValueError("inversion: need P probabilities and positions")
  _check(_lib.tspmi_inversion(_ptr(routes), P, n, _ptr(probs), _ptr(pos1), _ptr(pos2), float(rate), _stream(torch)),
         "tspmi_inversion")


def ox(parents, cut_lo, cut_hi, out) -> None:
  torch = require_device()
  _need(parents, torch.int32, "parents", 2)
  _need(out, torch.int32, "out", 2)
  P, n = parents.shape
  if out.shape != parents.shape:
    raise ValueError("ox: out shape != parents shape")
  if P >= 2:
    _need(cut_lo, torch.int32, "cut_lo", 1)
    _need(cut_hi, torch.int32, "cut_hi", 1)
    if cut_lo.shape[0] < P // 2 or cut_hi.shape[0] < P // 2:
      raise ValueError("ox: need P // 2 cut pairs")
  _check(_lib.tspmi_ox(_ptr(parents), P, n, _ptr(cut_lo), _ptr(cut_hi), _ptr(out), _stream(torch)), "tspmi_ox")


def swap(routes, probs, pos1, pos2, rate: float) -> None:
  torch = require_device()
  _need(routes, torch.int32, "routes", 2)
  _need(probs, torch.float64, "probs", 1)
  _need(pos1, torch.int32, "pos1", 1)
  _need(pos2, torch.int32, "pos2", 1)
  P, n = routes.shape
  if min(probs.shape[0], pos1.shape[0], pos2.shape[0]) < P:
    raise ValueError("swap: need P probabilities and positions")
  _check(_lib.tspmi_swap(_ptr(routes), P, n, _ptr(probs), _ptr(pos1), _ptr(pos2), float(rate), _stream(torch)),
         "tspmi_swap")


def generation(pop, fitness, D, tour_cand, cut_lo, cut_hi, probs, pos1, pos2, rate, offspring, off_fit) -> None:
  torch = require_device()
  P, n = pop.shape
  for t, dt, nm in ((pop, torch.int32, "pop"), (fitness, torch.float64, "fitness"), (D, torch.int32, "D"),
                    (tour_cand, torch.int32, "tour_cand"), (cut_lo, torch.int32, "cut_lo"),
                    (cut_hi, torch.int32, "cut_hi"), (probs, torch.float64, "probs"), (pos1, torch.int32, "pos1"),
                    (pos2, torch.int32, "pos2"), (offspring, torch.int32, "offspring"),
                    (off_fit, torch.float64, "off_fit")):
    _need(t, dt, nm)
  if tour_cand.shape[0] != P or offspring.shape != (P, n) or off_fit.shape[0] != P:
    raise ValueError("generation: shape mismatch")
  _check(_lib.tspmi_generation(_ptr(pop), _ptr(fitness), P, n, _ptr(D), D.shape[1], _ptr(tour_cand),
                               tour_cand.shape[1], _ptr(cut_lo), _ptr(cut_hi), _ptr(probs), _ptr(pos1), _ptr(pos2),
                               float(rate), _ptr(offspring), _ptr(off_fit), _stream(torch)), "tspmi_generation")


_M64 = (1 << 64) - 1


def draw_generations_pcg64(state: dict, G: int, P: int, n: int, k: int, record_states: bool = False):
  """Host draws of ``G`` generations from a PCG64 generator state (``bit_generator.state``), as
  tspmi_draw_generations makes them; no device work.  Returns (new state, cand (G, P, k),
  cut_lo, cut_hi (G, max(1, P // 2)), probs, pos1, pos2 (G, P), states (G, 6) or None)."""
  load_library()
  inner = state["state"]
  words = np.array([inner["state"] >> 64, inner["state"] & _M64, inner["inc"] >> 64, inner["inc"] & _M64,
                    int(state["has_uint32"]), int(state["uinteger"])], dtype=np.uint64)
  half = max(1, P // 2)
  cand = np.empty((G, P, k), np.int32)
  lo, hi = np.empty((G, half), np.int32), np.empty((G, half), np.int32)
  probs = np.empty((G, P), np.float64)
  p1, p2 = np.empty((G, P), np.int32), np.empty((G, P), np.int32)
  states = np.empty((G, 6), np.uint64) if record_states else None
  _check(_lib.tspmi_draw_generations(words.ctypes.data, G, P, n, k, cand.ctypes.data, lo.ctypes.data, hi.ctypes.data,
                                     probs.ctypes.data, p1.ctypes.data, p2.ctypes.data,
                                     None if states is None else states.ctypes.data), "tspmi_draw_generations")
  return pcg64_state(words, state), cand, lo, hi, probs, p1, p2, states


def pcg64_state(words, like: dict) -> dict:
  """The ``bit_generator.state`` dict of the six words {state_hi, state_lo, inc_hi, inc_lo,
  has_uint32, uinteger}."""
  w = [int(x) for x in words]
  return {"bit_generator": like["bit_generator"], "state": {"state": (w[0] << 64) | w[1], "inc": (w[2] << 64) | w[3]},
          "has_uint32": w[4], "uinteger": w[5]}


def rank_select(keys, keep: int, order) -> None:
  torch = require_device()
  _need(keys, torch.float64, "keys", 1)
  _need(order, torch.int32, "order", 1)
  if order.shape[0] < keep:
    raise ValueError("rank_select: order shorter than keep")
  _check(_lib.tspmi_rank_select(_ptr(keys), keys.shape[0], int(keep), _ptr(order), _stream(torch)), "tspmi_rank_select")


def survive(keys, keep: int, A, B, out, out_fit=None, order=None) -> None:
  """(mu + lambda) survival in one launch: ``out`` (keep, n) receives the rows of ``A`` then ``B``
  (``keys`` holds their fitness in that order) in stable rank order, ``out_fit`` their keys and
  ``order`` their source indices; as :func:`rank_select` followed by :func:`gather_rows`."""
  torch = require_device()
  _need(keys, torch.float64, "keys", 1)
  _need(A, torch.int32, "A", 2)
  _need(out, torch.int32, "out", 2)
  n = A.shape[1]
  M, P0 = keys.shape[0], A.shape[0]
  if B is not None:
    _need(B, torch.int32, "B", 2)
    if B.shape[1] != n or P0 + B.shape[0] < M:
      raise ValueError("survive: B does not match A / keys")
  elif P0 < M:
    raise ValueError("survive: keys beyond A's rows need B")
  if out.shape[0] < keep or out.shape[1] != n:
    raise ValueError("survive: out is not (keep, n)")
  if out_fit is not None:
    _need(out_fit, torch.float64, "out_fit", 1)
  if order is not None:
    _need(order, torch.int32, "order", 1)
  _check(_lib.tspmi_survive(_ptr(keys), M, int(keep), _ptr(A), P0, _ptr(B), n, _ptr(out), _ptr(out_fit), _ptr(order),
                            _stream(torch)), "tspmi_survive")


def gather_rows(A, fitA, B, fitB, src, out, out_fit=None) -> None:
  torch = require_device()
  _need(A, torch.int32, "A", 2)
  _need(src, torch.int32, "src", 1)
  _need(out, torch.int32, "out", 2)
  n = A.shape[1]
  _check(_lib.tspmi_gather_rows(_ptr(A), _ptr(fitA), A.shape[0], _ptr(B), _ptr(fitB), n, _ptr(src), src.shape[0],
                                _ptr(out), _ptr(out_fit), _stream(torch)), "tspmi_gather_rows")


def scatter_rows(src, src_fit, idx, dst, dst_fit=None) -> None:
  torch = require_device()
  _need(src, torch.int32, "src", 2)
  _need(idx, torch.int32, "idx", 1)
  _need(dst, torch.int32, "dst", 2)
  _check(_lib.tspmi_scatter_rows(_ptr(src), _ptr(src_fit), _ptr(idx), idx.shape[0], src.shape[1], _ptr(dst),
                                 _ptr(dst_fit), _stream(torch)), "tspmi_scatter_rows")


def random_init(seeds, customers, out) -> None:
  torch = require_device()
  _need(seeds, torch.int64, "seeds", 1)
  _need(customers, torch.int32, "customers", 1)
  _need(out, torch.int32, "out", 2)
  _check(_lib.tspmi_random_init(_ptr(seeds), out.shape[0], _ptr(customers), customers.shape[0], _ptr(out),
                                _stream(torch)), "tspmi_random_init")


def nn_init(seeds, customers, cand, work, out) -> None:
  torch = require_device()
  _need(seeds, torch.int64, "seeds", 1)
  _need(customers, torch.int32, "customers", 1)
  _need(out, torch.int32, "out", 2)
  _need(work, torch.int32, "work")
  k = 0 if cand is None else cand.shape[1]
  if work.numel() < 2 * out.shape[0] * out.shape[1]:
    raise ValueError("nn_init: work needs 2 * P * n int32")
  _check(_lib.tspmi_nn_init(_ptr(seeds), out.shape[0], _ptr(customers), customers.shape[0], _ptr(cand), k, _ptr(work),
                            _ptr(out), _stream(torch)), "tspmi_nn_init")


#: Largest n of tspmi_two_opt_full_euc2d (the diagonal scan's 8 slots of 256 lanes).
FULL_EUC2D_MAX_N = 2050


def full_euc2d_ws_bytes(n: int) -> int:
  """Workspace bytes of :func:`two_opt_full_euc2d` with ``ws`` (the renumbered instance)."""
  return int(load_library().tspmi_two_opt_full_euc2d_ws_bytes(int(n)))


def two_opt_full_euc2d(tours, D, xy, max_iter: int = 10, stats=None, ws=None) -> None:
  """Full-neighbourhood best-improvement 2-opt on an EUC_2D instance: distances computed from
  ``xy`` on most rows and gathered from the matrix ``D`` built from it on the others (n <= 2050).
  With ``ws`` (a device byte tensor of at least :func:`full_euc2d_ws_bytes` bytes), the instance
  is renumbered in Hilbert-curve order for the scan (tspmi_two_opt_full_euc2d_ordered: the same
  moves, tours and counters, fewer texture-addresser cycles per gather)."""
  torch = require_device()
  _need(tours, torch.int32, "tours", 2)
  _need(D, torch.int32, "D", 2)
  _need(xy, torch.float64, "xy", 2)
  if stats is not None:
    _need(stats, torch.int64, "stats")
  P, n = tours.shape
  if xy.shape != (n, 2) or D.shape[0] < n or D.shape[1] < n:
    raise ValueError(f"two_opt_full_euc2d: xy {tuple(xy.shape)} / D {tuple(D.shape)} do not match n={n}")
  if ws is not None:
    _need(ws, torch.uint8, "ws", 1)
    _check(_lib.tspmi_two_opt_full_euc2d_ordered(_ptr(tours), P, n, _ptr(D), D.shape[1], _ptr(xy), int(max_iter),
                                                 _ptr(stats), _ptr(ws), ws.numel(), _stream(torch)),
           "tspmi_two_opt_full_euc2d_ordered")
    return
  _check(_lib.tspmi_two_opt_full_euc2d(_ptr(tours), P, n, _ptr(D), D.shape[1], _ptr(xy), int(max_iter), _ptr(stats),
                                  _stream(torch)), "tspmi_two_opt_full_euc2d")


def two_opt_euc2d(tours, xy, max_iter: int = 10, stats=None) -> None:
  """Full-neighbourhood best-improvement 2-opt with distances computed from EUC_2D coordinates."""
  torch = require_device()
  _need(tours, torch.int32, "tours", 2)
  _need(xy, torch.float64, "xy", 2)
  if stats is not None:
    _need(stats, torch.int64, "stats")
  P, n = tours.shape
  if xy.shape != (n, 2):
    raise ValueError(f"two_opt_euc2d: xy {tuple(xy.shape)} != ({n}, 2)")
  _check(_lib.tspmi_two_opt_euc2d(_ptr(tours), P, n, _ptr(xy), int(max_iter), _ptr(stats), _stream(torch)),
         "tspmi_two_opt_euc2d")


def packed_euc2d_ws_bytes(n: int, k: int) -> int:
  """Workspace bytes of the renumbered candidate scan (``ws`` of :func:`two_opt_packed_euc2d`)."""
  return int(load_library().tspmi_two_opt_packed_euc2d_ws_bytes(int(n), int(k)))


def _ordered_call(tours, P, n, xy, table, k, max_iter, stats, cost, pass_log, log_moves, ws, torch) -> None:
  _need(ws, torch.uint8, "ws", 1)
  _check(_lib.tspmi_two_opt_packed_euc2d_i32_ordered(_ptr(tours), P, n, _ptr(xy), _ptr(table), int(k), int(max_iter),
                                                     _ptr(stats), _ptr(cost), _ptr(pass_log), int(log_moves),
                                                     _ptr(ws), ws.numel(), _stream(torch)),
         "tspmi_two_opt_packed_euc2d_i32_ordered")


def two_opt_packed_euc2d(tours, xy, table, k: int, max_iter: int = 10, stats=None, cost=None, ws=None) -> None:
  """Candidate-list best-improvement 2-opt on a packed table, distances from EUC_2D coordinates:
  float64 ``xy``, or int32 ``xy`` (integer coordinates spanning at most 2^20 per axis, as
  ``ProblemContext.get_gpu_int_coordinates`` gives them) for the integer distance form; same
  results.  ``cost`` as for :func:`two_opt_packed` (needs n >= 4 and max_iter != 0).  With int32
  ``xy`` and ``ws`` (a device byte tensor of :func:`packed_euc2d_ws_bytes` bytes) the scan runs on
  the instance renumbered in Hilbert order (same moves and tours)."""
  torch = require_device()
  _need(tours, torch.int32, "tours", 2)
  int_xy = isinstance(xy, torch.Tensor) and xy.dtype == torch.int32
  _need(xy, torch.int32 if int_xy else torch.float64, "xy", 2)
  _need(table, torch.int32, "table", 3)
  if stats is not None:
    _need(stats, torch.int64, "stats")
  P, n = tours.shape
  if xy.shape != (n, 2) or table.shape[0] != n or table.shape[1] != candidate_lanes(k):
    raise ValueError("two_opt_packed_euc2d: xy / table do not match n / k")
  if cost is not None:
    _need(cost, torch.float64, "cost", 1)
    if cost.shape[0] < P:
      raise ValueError("two_opt_packed_euc2d: cost shorter than the number of tours")
  if int_xy and ws is not None:
    _ordered_call(tours, P, n, xy, table, k, max_iter, stats, cost, None, 0, ws, torch)
  elif int_xy:
    _check(_lib.tspmi_two_opt_packed_euc2d_i32(_ptr(tours), P, n, _ptr(xy), _ptr(table), int(k), int(max_iter),
                                               _ptr(stats), _ptr(cost), None, 0, _stream(torch)),
           "tspmi_two_opt_packed_euc2d_i32")
  else:
    _check(_lib.tspmi_two_opt_packed_euc2d_cost(_ptr(tours), P, n, _ptr(xy), _ptr(table), int(k), int(max_iter),
                                                _ptr(stats), _ptr(cost), _stream(torch)),
           "tspmi_two_opt_packed_euc2d_cost")


def two_opt_packed_euc2d_log(tours, xy, table, k: int, max_iter: int, pass_log, stats=None, cost=None,
                             ws=None) -> None:
  """:func:`two_opt_packed_euc2d` with a per-move progress log: ``pass_log`` int64 (P, M + 1, 2)
  receives each tour's {cost, wall clock} at the start and after each of its first M moves (with
  ``ws``, of the scan on the renumbered instance: the clock starts after the renumbering)."""
  torch = require_device()
  _need(pass_log, torch.int64, "pass_log", 3)
  P, n = tours.shape
  if pass_log.shape[0] != P or pass_log.shape[2] != 2 or pass_log.shape[1] < 1:
    raise ValueError(f"two_opt_packed_euc2d_log: pass_log {tuple(pass_log.shape)} is not (P, M + 1, 2)")
  _need(tours, torch.int32, "tours", 2)
  int_xy = isinstance(xy, torch.Tensor) and xy.dtype == torch.int32
  _need(xy, torch.int32 if int_xy else torch.float64, "xy", 2)
  _need(table, torch.int32, "table", 3)
  if stats is not None:
    _need(stats, torch.int64, "stats")
  if xy.shape != (n, 2) or table.shape[0] != n or table.shape[1] != candidate_lanes(k):
    raise ValueError("two_opt_packed_euc2d_log: xy / table do not match n / k")
  if cost is not None:
    _need(cost, torch.float64, "cost", 1)
  if int_xy and ws is not None:
    _ordered_call(tours, P, n, xy, table, k, max_iter, stats, cost, pass_log, pass_log.shape[1] - 1, ws, torch)
    return
  fn = _lib.tspmi_two_opt_packed_euc2d_i32 if int_xy else _lib.tspmi_two_opt_packed_euc2d_log
  _check(fn(_ptr(tours), P, n, _ptr(xy), _ptr(table), int(k), int(max_iter), _ptr(stats), _ptr(cost), _ptr(pass_log),
            pass_log.shape[1] - 1, _stream(torch)),
         "tspmi_two_opt_packed_euc2d_i32" if int_xy else "tspmi_two_opt_packed_euc2d_log")


#: Largest n of the LDS-resident kernels: candidate scans keep (3n + 1) int32 in LDS next to 1 KiB
#: of reductions; the full-neighbourhood scans reach n = 16,386 (16 slots of 1024 lanes).
CAND_LDS_MAX_N = (160 * 1024 - 1024 - 4) // 12
FULL_LDS_MAX_N = 16 * 1024 + 2


def two_opt_large(tours, D=None, xy=None, table=None, k: int = 0, max_iter: int = 10, stats=None, cost=None,
                  ws=None) -> None:
  """Best-improvement 2-opt with the tour arrays in a global workspace (any n whose scan indices
  fit 32 bits): candidate scan on a packed ``table`` with ``k`` candidates, or the full
  neighbourhood when ``table`` is None; distances from ``xy`` (EUC_2D) if given, else from ``D``.
  ``ws`` (int32, P * (3n + 1)) is allocated when not given."""
  torch = require_device()
  _need(tours, torch.int32, "tours", 2)
  P, n = tours.shape
  if xy is None and D is None:
    raise ValueError("two_opt_large: needs D or xy")
  if xy is not None:
    _need(xy, torch.float64, "xy", 2)
    if xy.shape != (n, 2):
      raise ValueError(f"two_opt_large: xy {tuple(xy.shape)} != ({n}, 2)")
  if D is not None:
    _need(D, torch.int32, "D", 2)
  if table is not None:
    _need(table, torch.int32, "table", 3)
    if table.shape[0] != n or table.shape[1] != candidate_lanes(k):
      raise ValueError("two_opt_large: table does not match n / k")
  if stats is not None:
    _need(stats, torch.int64, "stats")
  if cost is not None:
    _need(cost, torch.float64, "cost", 1)
    if cost.shape[0] < P:
      raise ValueError("two_opt_large: cost shorter than the number of tours")
  need = P * (3 * n + 1)
  if ws is None:
    ws = torch.empty(max(1, need), dtype=torch.int32, device="cuda")
  else:
    _need(ws, torch.int32, "ws")
    if ws.numel() < need:
      raise ValueError(f"two_opt_large: ws needs {need} int32")
  _check(_lib.tspmi_two_opt_large(_ptr(tours), P, n, _ptr(D), 0 if D is None else D.shape[1], _ptr(xy), _ptr(table),
                                  int(k), int(max_iter), _ptr(stats), _ptr(cost), _ptr(ws), _stream(torch)),
         "tspmi_two_opt_large")


def wall_clock_khz() -> int:
  """Rate of the device's constant wall clock in kHz."""
  require_device()
  v = ctypes.c_int(0)
  _check(_lib.tspmi_wall_clock_khz(ctypes.addressof(v)), "tspmi_wall_clock_khz")
  return int(v.value)


def nn_greedy_init(starts, D, out) -> None:
  torch = require_device()
  _need(starts, torch.int32, "starts", 1)
  _need(D, torch.int32, "D", 2)
  _need(out, torch.int32, "out", 2)
  P, n = out.shape
  if starts.shape[0] != P or D.shape[0] < n:
    raise ValueError("nn_greedy_init: shape mismatch")
  _check(_lib.tspmi_nn_greedy_init(_ptr(starts), P, n, _ptr(D), D.shape[1], _ptr(out), _stream(torch)),
         "tspmi_nn_greedy_init")


# ---- uneven multi-instance batch (raw pointers; shapes are validated by the batch engine) -------


def batch_call(name: str, *args) -> None:
  """Call a ``tspmi_batch_*`` entry point with device tensors / ints / floats, on the current stream."""
  torch = require_device()
  conv = []
  for a in args:
    if isinstance(a, torch.Tensor):
      if not a.is_cuda or not a.is_contiguous():
        raise ValueError(f"{name}: tensor arguments must be contiguous device tensors")
      conv.append(int(a.data_ptr()))
    else:
      conv.append(a)
  _check(getattr(_lib, name)(*conv, _stream(torch)), name)


# ---- learned candidate scorer -------------------------------------------------------------------


def scorer_param_count() -> int:
  return int(load_library().tspmi_scorer_param_count())


def mlp_work_sizes(S: int) -> tuple[int, int]:
  load_library()
  w = ctypes.c_long()
  p = ctypes.c_long()
  _check(_lib.tspmi_mlp_work_sizes(int(S), ctypes.byref(w), ctypes.byref(p)), "tspmi_mlp_work_sizes")
  return int(w.value), int(p.value)


def scorer_features(xy, nn32, D, X) -> None:
  torch = require_device()
  _need(xy, torch.float64, "xy", 2)
  _need(nn32, torch.int32, "nn32", 2)
  _need(D, torch.int32, "D", 2)
  _need(X, torch.int16, "X", 2)
  n = xy.shape[0]
  if nn32.shape != (n, 32) or X.shape != (n * 32, 16):
    raise ValueError("scorer_features: shape mismatch")
  _check(_lib.tspmi_scorer_features(_ptr(xy), n, _ptr(nn32), _ptr(D), D.shape[1], _ptr(X), _stream(torch)),
         "tspmi_scorer_features")


def scorer_labels(tour, nn32, work, y) -> None:
  torch = require_device()
  _need(tour, torch.int32, "tour", 1)
  _need(nn32, torch.int32, "nn32", 2)
  _need(work, torch.int32, "work", 1)
  _need(y, torch.float32, "y", 1)
  n = tour.shape[0]
  if nn32.shape != (n, 32) or y.shape[0] != n * 32 or work.shape[0] < n:
    raise ValueError("scorer_labels: shape mismatch")
  _check(_lib.tspmi_scorer_labels(_ptr(tour), n, _ptr(nn32), _ptr(work), _ptr(y), _stream(torch)), "tspmi_scorer_labels")


def mlp_forward(X, wbf, z) -> None:
  torch = require_device()
  _need(X, torch.int16, "X", 2)
  _need(wbf, torch.int16, "wbf", 1)
  _need(z, torch.float32, "z", 1)
  if X.shape[1] != 16 or z.shape[0] < X.shape[0]:
    raise ValueError("mlp_forward: shape mismatch")
  _check(_lib.tspmi_mlp_forward(_ptr(X), X.shape[0], _ptr(wbf), _ptr(z), _stream(torch)), "tspmi_mlp_forward")


def mlp_grad(X, y, wbf, work, partials, grad) -> None:
  torch = require_device()
  _need(X, torch.int16, "X", 2)
  _need(y, torch.float32, "y", 1)
  _need(wbf, torch.int16, "wbf", 1)
  _need(work, torch.int16, "work", 1)
  _need(partials, torch.float32, "partials", 1)
  _need(grad, torch.float32, "grad", 1)
  S = X.shape[0]
  w, p = mlp_work_sizes(S)
  if work.numel() < w or partials.numel() < p or grad.numel() < scorer_param_count() + 1 or y.shape[0] < S:
    raise ValueError("mlp_grad: buffers too small")
  _check(_lib.tspmi_mlp_grad(_ptr(X), _ptr(y), S, _ptr(wbf), _ptr(work), _ptr(partials), _ptr(grad), _stream(torch)),
         "tspmi_mlp_grad")


def adam(theta, m, v, grad, step: int, lr: float, beta1: float, beta2: float, eps: float, weight_decay: float,
         wbf) -> None:
  torch = require_device()
  for t, nm in ((theta, "theta"), (m, "m"), (v, "v"), (grad, "grad")):
    _need(t, torch.float32, nm, 1)
  _need(wbf, torch.int16, "wbf", 1)
  _check(_lib.tspmi_adam(_ptr(theta), _ptr(m), _ptr(v), _ptr(grad), int(step), float(lr), float(beta1), float(beta2),
                         float(eps), float(weight_decay), _ptr(wbf), _stream(torch)), "tspmi_adam")


def to_bf16(src, dst) -> None:
  torch = require_device()
  _need(src, torch.float32, "src", 1)
  _need(dst, torch.int16, "dst", 1)
  _check(_lib.tspmi_to_bf16(_ptr(src), src.shape[0], _ptr(dst), _stream(torch)), "tspmi_to_bf16")


def scorer_topk(z, nn32, k: int, out) -> None:
  torch = require_device()
  _need(z, torch.float32, "z", 1)
  _need(nn32, torch.int32, "nn32", 2)
  _need(out, torch.int32, "out", 2)
  n = nn32.shape[0]
  if out.shape != (n, k) or z.shape[0] < n * 32:
    raise ValueError("scorer_topk: shape mismatch")
  _check(_lib.tspmi_scorer_topk(_ptr(z), _ptr(nn32), n, int(k), _ptr(out), _stream(torch)), "tspmi_scorer_topk")
