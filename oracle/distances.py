"""Oracle: EUC_2D nint distance matrix and k-NN candidate lists (test infrastructure only).

Restates ``src/distances/matrix.py`` of the reference:

* ``compute_distance_matrix`` follows ``matrix.py:214-273`` with the EUC_2D branch of
  ``_distance_block`` (``matrix.py:84-86``): ``int64(sqrt(dx^2 + dy^2) + 0.5)``, cast
  to int32, diagonal forced to zero (``matrix.py:271-272``).  Rows are processed in
  blocks of 512 above n = 2048 exactly like ``matrix.py:260-267``; for EUC_2D the
  blocking does not change any value.
* ``compute_candidate_list`` follows ``matrix.py:140-211``: self masked with the int64
  maximum, ``argpartition`` then ``lexsort((columns, distances))`` so ties go to the
  smaller node index — except at the k-th boundary, where ``argpartition`` decides.
"""

from __future__ import annotations

import numpy as np

CHUNK_THRESHOLD = 2048  # matrix.py:35
CHUNK_SIZE = 512  # matrix.py:38


def _euc2d_block(rows: np.ndarray, coords: np.ndarray) -> np.ndarray:
  """EUC_2D block, ``matrix.py:82-86``."""
  diff = rows[:, np.newaxis, :] - coords[np.newaxis, :, :]
  squared = np.sum(diff * diff, axis=2)
  return (np.sqrt(squared) + 0.5).astype(np.int64)


def compute_distance_matrix(coords: np.ndarray) -> np.ndarray:
  """Dense (n, n) int32 EUC_2D matrix, ``matrix.py:214-273``."""
  coords = np.asarray(coords, dtype=np.float64)
  n = int(coords.shape[0])
  if n <= CHUNK_THRESHOLD:
    dist = _euc2d_block(coords, coords)
  else:
    dist = np.concatenate(
      [_euc2d_block(coords[s : s + CHUNK_SIZE], coords) for s in range(0, n, CHUNK_SIZE)], axis=0
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
