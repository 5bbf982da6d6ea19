"""Host logic of the uneven batch: the task list of the one-launch-per-generation form."""

from __future__ import annotations

import numpy as np
import pytest

from tsp_gpu_clean_amd.batch import TASK_CHILD, TASK_ELITE, TASK_SURVIVE, fused_task_list


def _decode(tasks):
  t = np.asarray(tasks, dtype=np.uint32)
  return t >> 30, (t >> 12) & 0x3FFFF, t & 0xFFF


@pytest.mark.parametrize("P,E,cpt", [(64, 16, 4), (15, 0, 2), (12, 12, 8), (18, 5, 4), (64, 80, 8)])
def test_fused_task_list_order_and_counts(P, E, cpt):
  ns = np.array([300, 57, 2000, 57, 4, 1200])
  kind, inst, idx = _decode(fused_task_list(ns, P, E, cpt))
  e = min(E, P)
  groups = -(-P // cpt)
  # kinds in dependency order: every child task, then every survive task, then the elite rows
  assert (np.diff(kind.astype(np.int64)) >= 0).all()
  assert (kind == TASK_CHILD).sum() == groups * ns.size
  assert (kind == TASK_SURVIVE).sum() == ns.size
  assert (kind == TASK_ELITE).sum() == e * ns.size
  lpt = np.argsort(-ns, kind="stable")
  for k, per in ((TASK_CHILD, groups), (TASK_SURVIVE, 1), (TASK_ELITE, e)):
    sel = kind == k
    # instances longest first (ties in index order), each with its tasks 0 .. per - 1
    np.testing.assert_array_equal(inst[sel], np.repeat(lpt, per))
    np.testing.assert_array_equal(idx[sel], np.tile(np.arange(per), ns.size))
  # every child row of every instance is in exactly one child task
  for i in range(ns.size):
    rows = np.concatenate([np.arange(g * cpt, min(P, (g + 1) * cpt)) for g in idx[(kind == TASK_CHILD) & (inst == i)]])
    np.testing.assert_array_equal(np.sort(rows), np.arange(P))


def test_fused_task_list_rejects_unencodable_batches():
  with pytest.raises(ValueError):
    fused_task_list(np.full(1 << 18, 10), 8, 2, 2)
  with pytest.raises(ValueError):
    fused_task_list(np.array([10]), 4096, 2, 2)
