"""Minimal TSP problem record (fields of the reference's Problem, src/data_models/problem.py:64-461).

Only what the hot path reads is kept: ``name``, ``dimension``, ``problem_type``, ``edge_type``,
``coordinates`` or an explicit ``distances`` matrix, and ``return_to_depot``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

VALID_PROBLEM_TYPES = ("TSP", "ATSP")
COORDINATE_EDGE_TYPES = ("EUC_2D", "CEIL_2D", "MAN_2D", "MAX_2D", "ATT", "GEO")
SUPPORTED_EDGE_TYPES = (*COORDINATE_EDGE_TYPES, "EXPLICIT")


@dataclass(frozen=True)
class Constraints:
  return_to_depot: bool = True
  capacity: int | None = None


@dataclass(frozen=True, eq=False)
class Problem:
  name: str
  dimension: int
  problem_type: str = "TSP"
  edge_type: str = "EUC_2D"
  coordinates: np.ndarray | None = None
  distances: np.ndarray | None = None
  capacity: int | None = None
  demands: np.ndarray | None = None
  constraints: Constraints = field(default_factory=Constraints)

  def __post_init__(self) -> None:
    if self.problem_type not in VALID_PROBLEM_TYPES:
      raise ValueError(f"problem_type must be one of {VALID_PROBLEM_TYPES}, got {self.problem_type!r}")
    if self.edge_type not in SUPPORTED_EDGE_TYPES:
      raise ValueError(f"edge_type must be one of {SUPPORTED_EDGE_TYPES}, got {self.edge_type!r}")
    if self.dimension < 1:
      raise ValueError("dimension must be >= 1")
    if self.coordinates is None and self.distances is None:
      raise ValueError(f"Problem '{self.name}' has no distances or coordinates")
    if self.edge_type == "EXPLICIT" and self.distances is None:
      raise ValueError(f"Problem '{self.name}' is EXPLICIT but carries no distance matrix")
    if self.coordinates is not None and np.asarray(self.coordinates).shape != (self.dimension, 2):
      raise ValueError(f"coordinates must have shape ({self.dimension}, 2)")
    if self.distances is not None and np.asarray(self.distances).shape != (self.dimension, self.dimension):
      raise ValueError(f"distances must have shape ({self.dimension}, {self.dimension})")

  @property
  def is_vrp(self) -> bool:
    return False

  @property
  def is_symmetric(self) -> bool:
    return self.problem_type != "ATSP"

  @classmethod
  def from_coordinates(cls, name: str, coords, **kw) -> "Problem":
    xy = np.ascontiguousarray(np.asarray(coords, dtype=np.float64))
    return cls(name=name, dimension=int(xy.shape[0]), coordinates=xy, **kw)


_EXPLICIT_FORMATS = ("FULL_MATRIX", "UPPER_ROW", "LOWER_ROW", "UPPER_DIAG_ROW", "LOWER_DIAG_ROW", "UPPER_COL",
                     "LOWER_COL", "UPPER_DIAG_COL", "LOWER_DIAG_COL")


def explicit_matrix(values, n: int, fmt: str) -> np.ndarray:
  """(n, n) int32 matrix from an EDGE_WEIGHT_SECTION in TSPLIB95 format ``fmt`` (symmetric
  formats are mirrored; a column-wise triangle equals the opposite row-wise triangle)."""
  fmt = {"UPPER_COL": "LOWER_ROW", "LOWER_COL": "UPPER_ROW", "UPPER_DIAG_COL": "LOWER_DIAG_ROW",
         "LOWER_DIAG_COL": "UPPER_DIAG_ROW"}.get(fmt, fmt)
  vals = np.asarray(values, dtype=np.int64)
  out = np.zeros((n, n), dtype=np.int64)
  if fmt == "FULL_MATRIX":
    if vals.size != n * n:
      raise ValueError(f"FULL_MATRIX needs {n * n} values, got {vals.size}")
    out[:] = vals.reshape(n, n)
  else:
    rows, cols = {"UPPER_ROW": np.triu_indices(n, 1), "LOWER_ROW": np.tril_indices(n, -1),
                  "UPPER_DIAG_ROW": np.triu_indices(n, 0), "LOWER_DIAG_ROW": np.tril_indices(n, 0)}[fmt]
    if vals.size != rows.size:
      raise ValueError(f"{fmt} needs {rows.size} values for n={n}, got {vals.size}")
    out[rows, cols] = vals
    out[cols, rows] = vals
  if (out < 0).any() or out.max(initial=0) > np.iinfo(np.int32).max:
    raise ValueError("explicit distances must lie in [0, 2^31)")
  return out.astype(np.int32)


def load_tsplib_file(path: str) -> Problem:
  """TSPLIB95 reader for ``TYPE: TSP`` instances: NODE_COORD_SECTION with EDGE_WEIGHT_TYPE EUC_2D,
  CEIL_2D, MAN_2D, MAX_2D, ATT or GEO, or an EXPLICIT EDGE_WEIGHT_SECTION in any of the nine
  matrix formats.  DISPLAY_DATA_SECTION is skipped.

  The reference delegates parsing to the third-party ``routing-data`` package
  (src/loaders/tsplib_reader.py:138-146), which is not available here; this is an own reader of
  the TSPLIB95 file format for the instance kinds the reference's fixtures hold.
  """
  header: dict[str, str] = {}
  coords: list[tuple[float, float]] = []
  weights: list[str] = []
  section = None
  with open(path, encoding="utf-8") as fh:
    for raw in fh:
      line = raw.strip()
      if not line:
        continue
      up = line.upper()
      if up.startswith("EOF"):
        break
      if up.startswith(("NODE_COORD_SECTION", "EDGE_WEIGHT_SECTION", "DISPLAY_DATA_SECTION", "TOUR_SECTION",
                        "DEMAND_SECTION", "DEPOT_SECTION", "FIXED_EDGES_SECTION")):
        section = up.split()[0].rstrip(":")
        continue
      if section is None or (":" in line and not line[0].isdigit() and not line[0] in "+-."):
        if ":" in line:
          key, _, val = line.partition(":")
          header[key.strip().upper()] = val.strip()
          section = None
        continue
      if section == "NODE_COORD_SECTION":
        parts = line.split()
        if len(parts) >= 3:
          coords.append((float(parts[1]), float(parts[2])))
      elif section == "EDGE_WEIGHT_SECTION":
        weights.extend(line.split())
  name = header.get("NAME", path)
  ptype = header.get("TYPE", "TSP").split()[0]
  edge = header.get("EDGE_WEIGHT_TYPE", "").split()[0] if header.get("EDGE_WEIGHT_TYPE") else ""
  dim = int(header.get("DIMENSION", len(coords)))
  if edge == "EXPLICIT":
    fmt = header.get("EDGE_WEIGHT_FORMAT", "FULL_MATRIX").split()[0]
    if fmt not in _EXPLICIT_FORMATS:
      raise ValueError(f"{path}: unsupported EDGE_WEIGHT_FORMAT {fmt!r}")
    d = explicit_matrix([int(float(w)) for w in weights], dim, fmt)
    return Problem(name=name, dimension=dim, problem_type=ptype, edge_type="EXPLICIT", distances=d)
  if edge not in COORDINATE_EDGE_TYPES:
    raise ValueError(f"{path}: unsupported EDGE_WEIGHT_TYPE {edge!r}; supported: {SUPPORTED_EDGE_TYPES}")
  if len(coords) != dim:
    raise ValueError(f"{path}: DIMENSION {dim} but {len(coords)} coordinates")
  return Problem.from_coordinates(name, np.asarray(coords), problem_type=ptype, edge_type=edge)
