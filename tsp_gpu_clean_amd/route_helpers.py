"""Route batch shape and residency helpers (reference src/utils/route_helpers.py:24-309).

A device array is a torch tensor on the HIP device; a host array is NumPy.  Strategies keep
residency: device in -> device out with no host synchronisation, host in -> host out.
"""

from __future__ import annotations

from typing import Any

import numpy as np


def is_gpu_array(arr: Any) -> bool:
  """True for a torch tensor that lives on the HIP device (route_helpers.py:24-27)."""
  try:
    import torch
  except Exception:  # pragma: no cover
    return False
  return isinstance(arr, torch.Tensor) and arr.is_cuda


def ensure_routes_batch(routes: Any) -> tuple[Any, bool]:
  """(N, n) view plus a was-1D flag (route_helpers.py:234-258)."""
  if not hasattr(routes, "ndim"):
    routes = np.asarray(routes, dtype=np.int32)
  if routes.ndim == 1:
    return routes.reshape(1, -1), True
  if routes.ndim == 2:
    return routes, False
  raise ValueError(f"Expected 1D route or 2D routes batch, got array with ndim={routes.ndim}")


def restore_routes_shape(batch: Any, was_1d: bool) -> Any:
  """Undo ensure_routes_batch (route_helpers.py:285-309)."""
  return batch[0] if was_1d else batch


def to_device_int32(batch: Any, copy: bool = True):
  """Return (device int32 contiguous tensor, input_on_gpu, uploaded_bytes).

  A device input is copied (strategies never alias or mutate their input, two_opt.py:840-841);
  a host input is uploaded once.
  """
  import torch

  if is_gpu_array(batch):
    t = batch.to(dtype=torch.int32)
    t = t.clone() if (copy and t.data_ptr() == batch.data_ptr()) else t
    return t.contiguous(), True, 0
  arr = np.ascontiguousarray(np.asarray(batch, dtype=np.int32))
  t = torch.from_numpy(arr).to(device="cuda", non_blocking=False)
  return t, False, int(arr.nbytes)


def to_device_float64(values: Any):
  import torch

  if is_gpu_array(values):
    return values.to(dtype=torch.float64).contiguous(), 0
  arr = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
  return torch.from_numpy(arr).to(device="cuda"), int(arr.nbytes)


def to_host(t: Any) -> np.ndarray:
  return t.detach().cpu().numpy()


def resolve_route_bounds(n: int, context: Any = None, route_sample: Any = None) -> tuple[int, int]:
  """Mutable slice [low, high) (route_helpers.py:30-71) for the TSP permutation case.

  Depot-bounded VRP routes are out of scope; a host sample whose endpoints are equal, or a
  context declaring closed routes, keeps both endpoints fixed exactly as the reference does.
  """
  if route_sample is not None and n > 2 and not is_gpu_array(route_sample):
    sample = np.asarray(route_sample)
    if sample.size == n and int(sample[0]) == int(sample[-1]):
      return 1, n - 1
  if context is not None:
    return (1, n - 1) if bool(getattr(context, "is_closed", False)) else (0, n)
  if route_sample is not None:
    if is_gpu_array(route_sample):
      closed = n > 2 and int(route_sample[0]) == int(route_sample[-1])
      return (1, n - 1) if closed else (0, n)
    return 0, n
  return 1, n - 1
