"""The scratch-free elite 2-opt entry point (tspmi_two_opt_elite_noalloc) against
tspmi_two_opt_elite with a caller workspace, called through the C ABI: two calls in a row, the
second on a smaller instance, each bit-identical in tours, costs, best index and counters."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import c_two_opt
from oracle.distances import compute_candidate_list, compute_distance_matrix

pytestmark = pytest.mark.gpu


def _case(torch, nat, n, P, k, seed):
  xy = np.random.default_rng(seed).integers(0, 1_000_001, (n, 2)).astype(np.float64)
  d = compute_distance_matrix(xy)
  cand = compute_candidate_list(d, k)
  rng = np.random.default_rng(seed + 1)
  tours = np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)
  D = torch.from_numpy(d).cuda()
  table = torch.empty((n, nat.candidate_lanes(k), 2), dtype=torch.int32, device="cuda")
  nat.candidate_table(torch.from_numpy(cand).cuda(), D, table)
  return d, cand, tours, D, table


def _with_workspace(torch, nat, tours, D, table, k, iters):
  P = tours.shape[0]
  t = torch.from_numpy(tours).cuda()
  stats = torch.zeros(3, dtype=torch.int64, device="cuda")
  cost = torch.empty(P, dtype=torch.float64, device="cuda")
  best = torch.full((1,), -1, dtype=torch.int32, device="cuda")
  sync = torch.zeros(nat.ELITE_SYNC_WORDS * P, dtype=torch.int64, device="cuda")
  nat.two_opt_elite(t, D, table, k, iters, stats, cost, sync, best)
  return t.cpu().numpy(), cost.cpu().numpy(), int(best.item()), stats.cpu().tolist()


def _no_alloc(torch, nat, tours, D, table, k, iters):
  """Direct ctypes call of the C entry point."""
  P, n = tours.shape
  t = torch.from_numpy(tours).cuda()
  stats = torch.zeros(3, dtype=torch.int64, device="cuda")
  cost = torch.empty(P, dtype=torch.float64, device="cuda")
  best = torch.full((1,), -1, dtype=torch.int32, device="cuda")
  failed = torch.full((1,), -7, dtype=torch.int32, device="cuda")
  lib = nat.load_library()
  rc = lib.tspmi_two_opt_elite_noalloc(t.data_ptr(), P, n, D.data_ptr(), D.shape[1], table.data_ptr(), k, iters,
                                       stats.data_ptr(), cost.data_ptr(), best.data_ptr(), failed.data_ptr(),
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
  assert rc == 0, lib.tspmi_last_error()
  return t.cpu().numpy(), cost.cpu().numpy(), int(best.item()), stats.cpu().tolist(), int(failed.item())


def test_elite_noalloc_equals_elite_with_workspace():
  import torch

  from tsp_gpu_clean_amd import _native as nat

  nat.require_device()
  for n, P, k, iters, seed in ((1000, 16, 20, 10, 11), (300, 6, 8, -1, 12), (1000, 128, 20, 4, 13)):
    print(f"elite noalloc n={n} P={P} k={k} iters={iters}", flush=True)
    d, cand, tours, D, table = _case(torch, nat, n, P, k, seed)
    want = _with_workspace(torch, nat, tours, D, table, k, iters)
    got = _no_alloc(torch, nat, tours, D, table, k, iters)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])
    assert got[2] == want[2] and got[3] == want[3] and got[4] == 0
    ref, _ = c_two_opt.improve_batch(tours[:2], d, cand, "best", iters)
    np.testing.assert_array_equal(got[0][:2], ref)


def test_elite_noalloc_rejects_more_tours_than_its_workspace():
  import torch

  from tsp_gpu_clean_amd import _native as nat

  nat.require_device()
  d, cand, tours, D, table = _case(torch, nat, 50, nat.ELITE_STATIC_TOURS + 1, 8, 3)
  with pytest.raises(ValueError, match="static workspace"):
    nat.two_opt_elite_noalloc(torch.from_numpy(tours).cuda(), D, table, 8, 5, None, None)
