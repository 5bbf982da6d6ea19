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
