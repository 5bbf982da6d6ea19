"""Minimal TSP problem record (fields of the reference's Problem, src/data_models/problem.py:64-461).

Only what the hot path reads is kept: ``name``, ``dimension``, ``problem_type``, ``edge_type``,
``coordinates`` or an explicit ``distances`` matrix, and ``return_to_depot``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

VALID_PROBLEM_TYPES = ("TSP", "ATSP")
SUPPORTED_EDGE_TYPES = ("EUC_2D", "EXPLICIT")


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


def load_tsplib_file(path: str) -> Problem:
  """Minimal TSPLIB reader for ``TYPE: TSP`` with ``EDGE_WEIGHT_TYPE: EUC_2D`` node coordinates.

  The reference delegates parsing to the third-party ``routing-data`` package
  (src/loaders/tsplib_reader.py:138-146), which is not available here; this reader covers the
  EUC_2D instances the hot path runs on.
  """
  header: dict[str, str] = {}
  coords: list[tuple[float, float]] = []
  in_coords = False
  with open(path, encoding="utf-8") as fh:
    for raw in fh:
      line = raw.strip()
      if not line:
        continue
      if line.upper().startswith("EOF"):
        break
      if line.upper().startswith("NODE_COORD_SECTION"):
        in_coords = True
        continue
      if in_coords:
        parts = line.split()
        if len(parts) >= 3:
          coords.append((float(parts[1]), float(parts[2])))
        continue
      if ":" in line:
        key, _, val = line.partition(":")
        header[key.strip().upper()] = val.strip()
  edge = header.get("EDGE_WEIGHT_TYPE", "")
  if edge != "EUC_2D":
    raise ValueError(f"{path}: only EUC_2D node-coordinate instances are supported, got {edge!r}")
  dim = int(header.get("DIMENSION", len(coords)))
  if len(coords) != dim:
    raise ValueError(f"{path}: DIMENSION {dim} but {len(coords)} coordinates")
  return Problem.from_coordinates(header.get("NAME", path), np.asarray(coords), problem_type=header.get("TYPE", "TSP"))
