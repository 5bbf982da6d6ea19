"""Oracle: TSPLIB distance matrices and k-NN candidate lists (test infrastructure only).

Restates ``src/distances/matrix.py`` of the reference:

* ``compute_distance_matrix`` follows ``matrix.py:214-273`` with the 2D branches of
  ``_distance_block`` (``matrix.py:77-131``): EUC_2D ``int64(sqrt(dx^2 + dy^2) + 0.5)``,
  CEIL_2D ``ceil(sqrt(.))``, MAN_2D ``int64(|dx| + |dy| + 0.5)``, MAX_2D
  ``int64(max(|dx|, |dy|) + 0.5)``, ATT (pseudo-Euclidean ``sqrt(s / 10)`` rounded up when
  nint falls below it) and GEO (DDD.MM to radians with PI = 3.141592, great circle on
  R = 6378.388 km, ``int(R * acos(.) + 1)``, pairwise.py:62-102); cast to int32, diagonal
  forced to zero (``matrix.py:271-272``).  Rows are processed in blocks of 512 above
  n = 2048 exactly like ``matrix.py:260-267``; the blocking does not change any value.
* ``compute_candidate_list`` follows ``matrix.py:140-211``: self masked with the int64
  maximum, ``argpartition`` then ``lexsort((columns, distances))`` so ties go to the
  smaller node index — except at the k-th boundary, where ``argpartition`` decides.
"""

from __future__ import annotations

import numpy as np

CHUNK_THRESHOLD = 2048  # matrix.py:35
CHUNK_SIZE = 512  # matrix.py:38


GEO_PI = 3.141592  # pairwise.py:62
EARTH_RADIUS_KM = 6378.388  # pairwise.py:65
METRICS = ("EUC_2D", "CEIL_2D", "MAN_2D", "MAX_2D", "ATT", "GEO")


def _geo_radians(c: np.ndarray) -> np.ndarray:
  """DDD.MM to radians, matrix.py:45-49."""
  deg = np.trunc(c)
  return GEO_PI * (deg + 5.0 * (c - deg) / 3.0) / 180.0


def _block(rows: np.ndarray, coords: np.ndarray, metric: str) -> np.ndarray:
  """One (len(rows), n) block, matrix.py:77-131."""
  if metric == "GEO":
    lat_i = _geo_radians(rows[:, 0])[:, np.newaxis]
    lon_i = _geo_radians(rows[:, 1])[:, np.newaxis]
    lat_j = _geo_radians(coords[:, 0])[np.newaxis, :]
    lon_j = _geo_radians(coords[:, 1])[np.newaxis, :]
    q1 = np.cos(lon_i - lon_j)
    q2 = np.cos(lat_i - lat_j)
    q3 = np.cos(lat_i + lat_j)
    arg = np.clip(0.5 * ((1.0 + q1) * q2 - (1.0 - q1) * q3), -1.0, 1.0)
    return (EARTH_RADIUS_KM * np.arccos(arg) + 1.0).astype(np.int64)
  diff = rows[:, np.newaxis, :] - coords[np.newaxis, :, :]
  if metric == "EUC_2D":
    return (np.sqrt(np.sum(diff * diff, axis=2)) + 0.5).astype(np.int64)
  if metric == "CEIL_2D":
    return np.ceil(np.sqrt(np.sum(diff * diff, axis=2))).astype(np.int64)
  if metric == "MAN_2D":
    return (np.sum(np.abs(diff), axis=2) + 0.5).astype(np.int64)
  if metric == "MAX_2D":
    return (np.max(np.abs(diff), axis=2) + 0.5).astype(np.int64)
  if metric == "ATT":
    r = np.sqrt(np.sum(diff * diff, axis=2) / 10.0)
    t = (r + 0.5).astype(np.int64)
    return np.where(t < r, t + 1, t)
  raise ValueError(f"Unsupported edge_weight_type: {metric!r}")


def _euc2d_block(rows: np.ndarray, coords: np.ndarray) -> np.ndarray:
  """EUC_2D block, ``matrix.py:82-86``."""
  return _block(rows, coords, "EUC_2D")


def compute_distance_matrix(coords: np.ndarray, metric: str = "EUC_2D") -> np.ndarray:
  """Dense (n, n) int32 matrix of a 2D TSPLIB metric, ``matrix.py:214-273``."""
  coords = np.asarray(coords, dtype=np.float64)
  n = int(coords.shape[0])
  if n <= CHUNK_THRESHOLD:
    dist = _block(coords, coords, metric)
  else:
    dist = np.concatenate(
      [_block(coords[s : s + CHUNK_SIZE], coords, metric) for s in range(0, n, CHUNK_SIZE)], axis=0
    )
  out = dist.astype(np.int32)
  np.fill_diagonal(out, 0)
  return out


def compute_candidate_list(dist: np.ndarray, k: int, chunk_size: int = CHUNK_SIZE) -> np.ndarray:
  """(n, min(k, n-1)) int32 nearest-neighbour table, ``matrix.py:140-211``."""
  n = int(dist.shape[0])
  if n <= 1 or k <= 0:
    return np.zeros((n, 0), dtype=np.int32)
  k_eff = min(k, n - 1)
  sentinel = int(np.iinfo(np.int64).max)
  result = np.empty((n, k_eff), dtype=np.int32)
  for start in range(0, n, max(1, chunk_size)):
    end = min(start + chunk_size, n)
    block = np.asarray(dist[start:end], dtype=np.int64).copy()
    rows = np.arange(end - start)
    block[rows, np.arange(start, end)] = sentinel
    cols = np.argpartition(block, k_eff, axis=1)[:, :k_eff]
    cdist = block[rows[:, np.newaxis], cols]
    order = np.lexsort((cols, cdist), axis=1)
    result[start:end, :] = cols[rows[:, np.newaxis], order]
  return result


def candidate_boundary_ties(dist: np.ndarray, k: int) -> int:
  """Count rows whose k-th and (k+1)-th nearest distances tie (``argpartition`` ambiguity).

  On such rows the reference's choice of which tied node enters ``N_k`` is
  implementation-defined; tests only compare candidate tables on rows without such
  a tie.
  """
  n = int(dist.shape[0])
  k_eff = min(k, n - 1)
  if k_eff >= n - 1 or k_eff <= 0:
    return 0
  d = np.asarray(dist, dtype=np.int64).copy()
  np.fill_diagonal(d, np.iinfo(np.int64).max)
  part = np.partition(d, k_eff, axis=1)
  kth = np.partition(d, k_eff - 1, axis=1)[:, k_eff - 1]
  return int(np.count_nonzero(kth == part[:, k_eff]))
