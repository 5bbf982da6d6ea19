"""Pins of the scorer oracle (oracle/mlp.py) on the CPU: the analytic float64 gradient against
central finite differences, the device-rounding gradient against the float64 one, bf16 rounding
against torch's bfloat16 cast, Adam against its closed form, and the edge features against the
invariants their definitions fix."""

import numpy as np
import pytest
import torch

from oracle import mlp
from oracle.distances import compute_candidate_list, compute_distance_matrix


def _theta(seed=0, scale=1.0):
  th = mlp.init_params(np.random.default_rng(seed)).astype(np.float64)
  rng = np.random.default_rng(seed + 1)
  th[mlp.OFF_B1 : mlp.OFF_W2] = rng.normal(0, 0.1, mlp.H)
  th[mlp.OFF_B2 : mlp.OFF_W3] = rng.normal(0, 0.1, mlp.H)
  th[mlp.OFF_B3] = 0.05
  return th * scale


def _batch(S=24, seed=3):
  rng = np.random.default_rng(seed)
  x = np.zeros((S, mlp.F))
  x[:, : mlp.NFEAT] = rng.normal(0, 1, (S, mlp.NFEAT))
  y = (rng.random(S) < 0.3).astype(np.float64)
  return x, y


def test_float64_gradient_matches_finite_differences():
  x, y = _batch()
  th = _theta()
  g = mlp.float64_grad(x, y, th)
  rng = np.random.default_rng(9)
  probes = [mlp.OFF_B3]
  for lo, hi in ((mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
                 (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3)):
    probes += [int(p) for p in rng.integers(lo, hi, 12)]
  eps = 1e-6
  for p in probes:
    if mlp.OFF_W1 <= p < mlp.OFF_B1 and p % mlp.F >= mlp.NFEAT:
      assert g[p] == 0.0
      continue
    tp, tm = th.copy(), th.copy()
    tp[p] += eps
    tm[p] -= eps
    fd = (mlp.float64_loss(x, y, tp) - mlp.float64_loss(x, y, tm)) / (2 * eps)
    assert abs(fd - g[p]) <= 1e-6 + 1e-4 * abs(fd), (p, fd, g[p])


def test_loss_is_bce_with_logits():
  # Single sample, closed form: z = b3 when every weight is zero.
  th = np.zeros(mlp.NPARAM)
  th[mlp.OFF_B3] = 0.7
  x = np.zeros((1, mlp.F))
  for yv in (0.0, 1.0):
    want = -(yv * np.log(1 / (1 + np.exp(-0.7))) + (1 - yv) * np.log(1 - 1 / (1 + np.exp(-0.7))))
    assert mlp.float64_loss(x, np.array([yv]), th) == pytest.approx(want, rel=1e-12)
    g = mlp.float64_grad(x, np.array([yv]), th)
    assert g[mlp.OFF_B3] == pytest.approx(1 / (1 + np.exp(-0.7)) - yv, rel=1e-12)


def test_rounded_gradient_tracks_float64_gradient():
  x, y = _batch(S=256)
  th = _theta(4)
  xb = mlp.bf16_bits(x.astype(np.float32))
  wb = mlp.bf16_bits(th.astype(np.float32))
  loss, g = mlp.loss_and_grad(xb, y, wb)
  g64 = mlp.float64_grad(mlp.from_bits(xb), y, mlp.from_bits(wb))
  assert loss == pytest.approx(mlp.float64_loss(mlp.from_bits(xb), y, mlp.from_bits(wb)), rel=2e-2)
  for lo, hi in ((mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
                 (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.NPARAM)):
    err = np.linalg.norm(g[lo:hi] - g64[lo:hi]) / max(np.linalg.norm(g64[lo:hi]), 1e-30)
    assert err < 2e-2, (lo, err)


def test_bf16_rounding_matches_torch_cast():
  rng = np.random.default_rng(0)
  vals = np.concatenate([rng.normal(0, 1, 4096).astype(np.float32),
                         np.array([1.0 + 2**-8, 1.0 + 3 * 2**-8, -1.0 - 2**-8, 0.0, -0.0, 1e-30, 3e38], np.float32)])
  want = torch.from_numpy(vals).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
  assert np.array_equal(mlp.bf16_bits(vals), want)
  assert mlp.bf16_round(np.float32(1.0 + 2**-8)) == 1.0  # tie to even
  assert mlp.bf16_round(np.float32(1.0 + 3 * 2**-8)) == 1.0 + 2**-6


def test_adam_closed_form():
  th0 = _theta(2).astype(np.float32)
  rng = np.random.default_rng(5)
  g = rng.normal(0, 1, mlp.NPARAM).astype(np.float32)
  pad = np.zeros(mlp.NPARAM, bool)
  pad[mlp.OFF_W1 : mlp.OFF_B1] = (np.arange(mlp.H * mlp.F) % mlp.F) >= mlp.NFEAT
  z = np.zeros(mlp.NPARAM, np.float32)
  lr = 1e-3
  th1, m1, v1 = mlp.adam(th0, z, z, g, 1, lr=lr)
  # bias-corrected first step moves every live parameter by lr * g / (|g| + eps)
  want = th0 - lr * g / (np.abs(g) + 1e-8)
  assert np.allclose(th1[~pad], want[~pad], rtol=0, atol=1e-6)
  assert (th1[pad] == 0).all()
  # a constant gradient keeps m_hat = g and v_hat = g^2, so step two moves by the same amount
  th2, _, _ = mlp.adam(th1, m1, v1, g, 2, lr=lr)
  assert np.allclose(th2[~pad], (th1 - lr * g / (np.abs(g) + 1e-8))[~pad], rtol=0, atol=2e-6)
  # decoupled from the bias terms: weight decay adds wd * theta to the weights' gradient only
  th3, _, _ = mlp.adam(th0, z, z, z, 1, lr=lr, wd=0.1)
  w = np.zeros(mlp.NPARAM, bool)
  w[mlp.OFF_W1 : mlp.OFF_B1] = True
  w[mlp.OFF_W2 : mlp.OFF_B2] = True
  w[mlp.OFF_W3 : mlp.OFF_B3] = True
  live = w & ~pad & (np.abs(th0) > 1e-3)
  assert np.allclose(th3[live], (th0 - lr * np.sign(th0))[live], atol=1e-6)
  assert np.array_equal(th3[~w], th0[~w])


def test_features_invariants():
  rng = np.random.default_rng(1)
  n = 120
  xy = rng.integers(0, 10_000, (n, 2)).astype(np.float64)
  D = compute_distance_matrix(xy)
  nn = compute_candidate_list(D, 32)
  X = mlp.features(xy, nn, D)
  u = np.repeat(np.arange(n), 32)
  r = np.tile(np.arange(32), n)
  v = nn.reshape(-1)
  assert np.allclose(X[:, 0] * 31, r)
  assert (X[:, 1] <= 1.0 + 1e-7).all()  # v lies within u's 32 nearest
  assert np.all(X[r < 8, 9] <= 1.0 + 1e-7)
  assert ((X[:, 10] >= 0) & (X[:, 10] <= 1.0 + 1e-6)).all()
  # mutuality agrees with the reverse rank and with membership of u in v's list
  member = np.array([u[t] in nn[v[t]] for t in range(n * 32)])
  assert np.array_equal(X[:, 6] == 1.0, member)
  assert np.all(X[member, 5] <= 1.0) and np.all(X[~member, 5] == 1.5)
  # the offset features rebuild the Euclidean distance within the rounding of D
  du32 = np.maximum(D[u, nn[u, 31]], 1)
  eu = np.hypot(X[:, 7], X[:, 8]) * du32
  assert np.all(np.abs(eu - D[u, v]) <= 0.5 + 1e-3 * D[u, v])
  assert (X[:, 12:] == 0).all()
