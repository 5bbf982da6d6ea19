"""Pins the distance / candidate-list oracle with the reference's known answers."""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle.distances import candidate_boundary_ties, compute_candidate_list, compute_distance_matrix
from tests.conftest import GOLDEN


def test_three_point_euc2d_known_answer():
  # reference tests/distances/test_unit_distance_matrix.py:29,97
  pts = np.array([[0.0, 0.0], [3.1, 4.0], [1.5, 2.5]])
  assert compute_distance_matrix(pts).tolist() == [[0, 5, 3], [5, 0, 2], [3, 2, 0]]


def test_berlin52_known_distances(berlin52_coords):
  # reference tests/distances/test_unit_distance_matrix.py:294-313
  d = compute_distance_matrix(berlin52_coords)
  assert d.shape == (52, 52) and d.dtype == np.int32
  assert d[0, 1] == 666 and d[0, 2] == 281 and d[1, 2] == 649
  assert np.array_equal(d, d.T) and not d.diagonal().any()


def test_nint_closed_form_on_integer_coordinates():
  rng = np.random.default_rng(42)
  xy = rng.integers(0, 1_000_001, (300, 2)).astype(np.float64)
  d = compute_distance_matrix(xy)
  ix = xy.astype(np.int64)
  s = ((ix[:, None, :] - ix[None, :, :]) ** 2).sum(axis=2)
  # exact integer nint: the integer r with (r - 1/2)^2 <= s < (r + 1/2)^2, i.e. 4s in [(2r-1)^2, (2r+1)^2)
  r = np.floor(np.sqrt(s.astype(np.float64))).astype(np.int64)
  r = np.where((2 * r + 1) ** 2 <= 4 * s, r + 1, r)
  np.fill_diagonal(r, 0)
  assert np.array_equal(d.astype(np.int64), r)


def test_chunked_path_matches_direct():
  rng = np.random.default_rng(3)
  xy = rng.integers(0, 1000, (2100, 2)).astype(np.float64)
  d = compute_distance_matrix(xy)
  sub = compute_distance_matrix(xy[:700])
  assert np.array_equal(d[:700, :700], sub)


def test_candidate_list_brute_force_without_boundary_ties():
  rng = np.random.default_rng(0)
  xy = rng.integers(0, 1_000_001, (400, 2)).astype(np.float64)
  d = compute_distance_matrix(xy)
  for k in (5, 20):
    assert candidate_boundary_ties(d, k) == 0
    got = compute_candidate_list(d, k)
    dd = d.astype(np.int64)
    np.fill_diagonal(dd, np.iinfo(np.int64).max)
    want = np.argsort(dd, axis=1, kind="stable")[:, :k]
    assert np.array_equal(got, want)


def test_candidate_list_clamps_and_empty():
  d = compute_distance_matrix(np.array([[0.0, 0.0], [1.0, 0.0], [5.0, 0.0]]))
  assert compute_candidate_list(d, 20).tolist() == [[1, 2], [0, 2], [1, 0]]
  assert compute_candidate_list(d[:1, :1], 3).shape == (1, 0)


# Known answers held by the reference's distance tests (test_unit_distance_matrix.py:93-125, 273-292):
# three hand-checked points per metric, and burma14's published GEO distances.
CARTESIAN_POINTS = np.array([[0.0, 0.0], [3.1, 4.0], [1.5, 2.5]])
GEO_POINTS = np.array([[16.47, 96.10], [16.47, 94.44], [20.09, 92.54]])


@pytest.mark.parametrize("metric,expected", [
  ("EUC_2D", [[0, 5, 3], [5, 0, 2], [3, 2, 0]]),
  ("CEIL_2D", [[0, 6, 3], [6, 0, 3], [3, 3, 0]]),
  ("MAN_2D", [[0, 7, 4], [7, 0, 3], [4, 3, 0]]),
  ("MAX_2D", [[0, 4, 3], [4, 0, 2], [3, 2, 0]]),
  ("ATT", [[0, 2, 1], [2, 0, 1], [1, 1, 0]]),
  ("GEO", [[0, 153, 510], [153, 0, 422], [510, 422, 0]]),
])
def test_metric_known_answers(metric, expected):
  pts = GEO_POINTS if metric == "GEO" else CARTESIAN_POINTS
  assert compute_distance_matrix(pts, metric).tolist() == expected


def test_metric_edge_cases():
  same = np.array([[1.5, 2.5], [1.5, 2.5]])
  for m in ("EUC_2D", "CEIL_2D", "MAN_2D", "MAX_2D", "ATT"):
    assert compute_distance_matrix(same, m).tolist() == [[0, 0], [0, 0]]
  assert compute_distance_matrix(np.array([[16.47, 96.10], [16.47, 96.10]]), "GEO").tolist() == [[0, 1], [1, 0]]
  for m in ("EUC_2D", "CEIL_2D", "MAN_2D", "MAX_2D", "ATT"):  # translation invariance
    a = compute_distance_matrix(np.array([[0.0, 0.0], [3.0, 4.0]]), m)
    b = compute_distance_matrix(np.array([[-7.0, -11.0], [-4.0, -7.0]]), m)
    assert a.tolist() == b.tolist()
  with pytest.raises(ValueError):
    compute_distance_matrix(CARTESIAN_POINTS, "EXPLICIT")


def test_burma14_geo_distances():
  from tsp_gpu_clean_amd.problem import load_tsplib_file

  prob = load_tsplib_file(os.path.join(GOLDEN, "burma14.tsp"))
  assert prob.edge_type == "GEO" and prob.dimension == 14
  d = compute_distance_matrix(prob.coordinates, "GEO")
  assert d[0, :3].tolist() == [0, 153, 510] and d[1, 2] == 422
  assert np.array_equal(d, d.T) and (np.diag(d) == 0).all()


def test_integer_nint_form_equals_fp64_nint():
  """The integer form of the EUC_2D distance used by the full-neighbourhood scan for integer
  instances (m = trunc of an FP32 sqrt(dx^2 + dy^2), then m + [t < 0] with t = m^2 + m - s
  formed modulo 2^32 from the 24-bit products) returns the oracle's FP64 nint for both estimates
  it can meet, nint - 1 and nint, and moves the wrong one (nint + 1 would stay: the estimate
  must never reach it); an FP32 estimate is within those two, with room to spare (the sqrt is
  never further than 0.3 from exact): on the nint-boundary offsets (both sides of each
  boundary) and random offsets up to the 2^20 span."""
  from tests.conftest import INT_SPAN, nint_boundary_offsets

  rng = np.random.default_rng(5)
  offs = np.concatenate([nint_boundary_offsets(), rng.integers(-INT_SPAN, INT_SPAN + 1, (4000, 2)),
                         [[INT_SPAN, INT_SPAN], [INT_SPAN, 0], [0, 0], [1, 0], [1, 1]]])
  pts = np.concatenate([[[0, 0]], offs]).astype(np.float64)
  want = compute_distance_matrix(pts)[0, 1:].astype(np.int64)
  dx, dy = offs[:, 0], offs[:, 1]
  s32 = (dx * dx + dy * dy) & 0xFFFFFFFF

  def correct(m, s=s32):
    t = ((m * m + m - s) & 0xFFFFFFFF).astype(np.int64)
    t = np.where(t >= 1 << 31, t - (1 << 32), t)
    return m + (t < 0)

  np.testing.assert_array_equal(correct(want), want)
  low = want > 0
  np.testing.assert_array_equal(correct(want[low] - 1, s32[low]), want[low])
  assert (correct(want + 1) == want + 1).all()  # an overshoot is not repaired
  fx, fy = dx.astype(np.float32), dy.astype(np.float32)
  r = np.sqrt(fx * fx + fy * fy).astype(np.float64)
  exact = np.sqrt((dx * dx + dy * dy).astype(np.float64))
  assert np.abs(r - exact).max() < 0.3
  est = r.astype(np.int64)
  assert ((est == want) | (est == want - 1)).all()
  np.testing.assert_array_equal(correct(est), want)
  assert (est != want).any()  # the correction does work on these pairs
