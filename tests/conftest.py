"""Shared pytest configuration: the ``gpu`` marker and small problem fixtures."""

from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
  config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libtspmi.so")
  config.addinivalue_line("markers", "slow: long-running CPU test")


def _gpu_available() -> bool:
  try:
    import torch

    return bool(torch.cuda.is_available())
  except Exception:
    return False


def pytest_collection_modifyitems(config, items):
  if _gpu_available():
    return
  skip = pytest.mark.skip(reason="no HIP device visible")
  for item in items:
    if "gpu" in item.keywords:
      item.add_marker(skip)


@pytest.fixture(scope="session")
def berlin52_coords() -> np.ndarray:
  return np.loadtxt(os.path.join(GOLDEN, "berlin52_coords.txt"), dtype=np.float64)


def rounded_matrix(n: int, seed: int) -> np.ndarray:
  """Symmetric EUC_2D-shaped int32 matrix from float coordinates in [0, 1000)^2."""
  rng = np.random.default_rng(seed)
  coords = rng.random((n, 2)) * 1000.0
  diff = coords[:, None, :] - coords[None, :, :]
  d = np.floor(np.sqrt((diff**2).sum(axis=2)) + 0.5).astype(np.int32)
  np.fill_diagonal(d, 0)
  return d


def lattice_matrix(n: int) -> np.ndarray:
  """Integer-lattice matrix: many tied distances, so tie-breaking is observable."""
  side = int(np.ceil(np.sqrt(n)))
  coords = np.array([[x, y] for x in range(side) for y in range(side)][:n], dtype=float)
  diff = coords[:, None, :] - coords[None, :, :]
  d = np.floor(np.sqrt((diff**2).sum(axis=2)) + 0.5).astype(np.int32)
  np.fill_diagonal(d, 0)
  return d


def nearest(dist: np.ndarray, k: int) -> np.ndarray:
  """N_k with self excluded, ascending distance, stable (lower index first)."""
  n = dist.shape[0]
  k_eff = min(k, n - 1)
  return np.asarray(np.argsort(dist, axis=1, kind="stable")[:, 1 : k_eff + 1], dtype=np.int32)


INT_SPAN = 1 << 20  # per-axis span up to which the integer EUC_2D form is exact


def nint_boundary_offsets() -> np.ndarray:
  """Integer offsets (dx, dy) whose squared length s sits at an nint boundary or next to it:
  (t^2 - c, t) gives s = m^2 + m + c with m = t^2 - c, the last s whose nint is m (c = 0), the
  first whose nint is m + 1 (c = 1) and their neighbours (c = -1, 2); t up to the 2^20 span, in
  all four sign quadrants and with the axes swapped."""
  ts = sorted({int(v) for v in np.geomspace(1, 1024, 40)} | {1, 2, 3, 1000, 1023, 1024})
  offs = []
  for t in ts:
    for c in (-1, 0, 1, 2):
      dx = t * t - c
      if dx < 0 or dx > INT_SPAN:
        continue
      for sx in (1, -1):
        for sy in (1, -1):
          offs += [(sx * dx, sy * t), (sy * t, sx * dx)]
  return np.array(offs, dtype=np.int64)


def span_limit_int_coords(n: int, seed: int) -> np.ndarray:
  """n integer points whose per-axis span is exactly 2^20: the corner (-2^19, -2^19), points at
  the nint-boundary offsets from it (folded into the box), the opposite corner, and uniform fill."""
  half = INT_SPAN // 2
  base = np.array([-half, -half], dtype=np.int64)
  offs = np.abs(nint_boundary_offsets())
  pts = [base, base + INT_SPAN] + [base + o for o in offs]
  rng = np.random.default_rng(seed)
  if len(pts) < n:
    pts += list(rng.integers(-half, half + 1, (n - len(pts), 2)))
  return np.array(pts[:n], dtype=np.float64)
