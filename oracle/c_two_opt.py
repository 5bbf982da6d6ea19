"""ctypes binding of ``two_opt_c.c`` (test infrastructure only).

The C restatement follows the same reference lines as ``oracle/two_opt.py`` for the
cyclic TSP topology; tests cross-check the two on random inputs before either is
trusted as the parity reference of the HIP kernels.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_two_opt.so")
_lib = None


def build() -> str:
  subprocess.run(["make", "-s", "-C", _HERE], check=True)
  return _LIB_PATH


def _load():
  global _lib
  if _lib is None:
    if not os.path.exists(_LIB_PATH):
      build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.oracle_two_opt.restype = ctypes.c_int
    lib.oracle_two_opt.argtypes = [
      ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_void_p,
    ]
    _lib = lib
  return _lib


def improve_batch(routes, dist, candidates=None, mode="best", max_iterations=10):
  """Return (improved routes, stats[deltas, moves, passes]); max_iterations < 0 = to optimum."""
  lib = _load()
  tours = np.array(routes, dtype=np.int32, copy=True, order="C")
  was_1d = tours.ndim == 1
  tours2 = tours.reshape(1, -1) if was_1d else tours
  d = np.ascontiguousarray(dist, dtype=np.int32)
  if candidates is not None:
    cand = np.ascontiguousarray(candidates, dtype=np.int32)
    k = int(cand.shape[1])
    cptr = cand.ctypes.data
  else:
    cand, k, cptr = None, 0, None
  stats = np.zeros(3, dtype=np.int64)
  rc = lib.oracle_two_opt(tours2.ctypes.data, tours2.shape[0], tours2.shape[1], d.ctypes.data,
                          d.shape[1], cptr, k, {"best": 0, "first": 1, "multi_move": 2}[mode], int(max_iterations),
                          stats.ctypes.data)
  if rc != 0:
    raise ValueError("oracle_two_opt rejected its arguments")
  return (tours2[0] if was_1d else tours2), stats
