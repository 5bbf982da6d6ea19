"""CPU oracle for the memetic-TSP hot path — TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy (and in plain C for the 2-opt scan, see
``two_opt_c.c``), the reference CPU algorithms of ``lucasgaldinos/tsp_gpu_clean``
that the HIP kernels of ``tsp_gpu_clean_amd`` must reproduce bit for bit.  Every
function cites the reference ``file:line`` it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it, and only as the parity reference or the timed CPU baseline.  The product
package never imports it and has no CPU fallback.

Pinning (what checks each part independently of its own formula):

* ``distances``: TSPLIB nint closed form, the reference's hand-computed 3-point
  EUC_2D matrix and the berlin52 known distances (tests/test_oracle_distances.py).
* ``two_opt``: the reference tests' nested-loop scalar specification (restated in
  tests/test_oracle_two_opt.py), pair-order and duplicate-candidate known answers,
  the per-move identity cost(after) = cost(before) + delta, and brute-force local
  optimality of the converged tours.
* ``operators``: hand-worked PMX/inversion/tournament examples and permutation
  invariants; the RNG draw order is fixed by the reference call sites.
* ``init``: SplitMix64 published constants (Steele, Lea & Flood 2014), a scalar
  Python cross-check of the vectorised Fisher–Yates, and permutation invariants.
* ``ga``: the reference's evolve() contract (history length, monotone best, stop
  reasons) plus a per-generation check against the operator oracles.

Parity unpinned: default-kind ``np.argsort`` tie order (elite choice at a fitness
tie on the 16th boundary) and ``np.argpartition`` boundary ties in the kNN table are
implementation-defined in NumPy; the oracle calls the same NumPy routines and the
tests assert the absence of such ties in their inputs.
"""
