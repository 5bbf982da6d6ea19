"""Oracle: the candidate-scorer MLP, its gradient, Adam and the edge features, in NumPy float64
(test infrastructure only).

There is no reference implementation of this component (the learned scorer extends the reference
with a new candidate method), so the oracle is a restatement of the published method it implements — a
12 -> 128 -> 128 -> 1 ReLU MLP trained with binary cross-entropy on logits and Adam (Kingma & Ba,
2015) — evaluated in float64 while rounding to bf16 exactly where the device kernels do:
inputs x, the weights, relu(h1) before layer 2, the layer-2 deltas before the W2^T product and
the weight-gradient products, relu(h1) in the weight-gradient product, and delta1 before the
W1 gradient.  Pinned independently by central finite differences on the float64 loss
(tests/test_oracle_mlp.py).

Parameter layout (shared with the device): W1 [128][16] (features 12..15 zero), b1 [128],
W2 [128][128], b2 [128], w3 [128], b3 [1].
"""

from __future__ import annotations

import numpy as np

H, F, NFEAT = 128, 16, 12
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1


def bf16_round(a) -> np.ndarray:
  """Round float32 values to bf16 (round to nearest even), returned as float64."""
  f = np.asarray(a, dtype=np.float32)
  u = f.view(np.uint32).astype(np.uint64)
  r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
  return r.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a) -> np.ndarray:
  f = np.asarray(a, dtype=np.float32)
  u = f.view(np.uint32).astype(np.uint64)
  return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bits(b) -> np.ndarray:
  return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(theta):
  t = np.asarray(theta, dtype=np.float64)
  return (t[OFF_W1:OFF_B1].reshape(H, F), t[OFF_B1:OFF_W2], t[OFF_W2:OFF_B2].reshape(H, H), t[OFF_B2:OFF_W3],
          t[OFF_W3:OFF_B3], t[OFF_B3])


def init_params(rng: np.random.Generator) -> np.ndarray:
  """He-normal weights, zero biases, zero padding features (float32)."""
  th = np.zeros(NPARAM, dtype=np.float32)
  w1 = rng.normal(0.0, np.sqrt(2.0 / NFEAT), (H, NFEAT))
  th[OFF_W1:OFF_B1].reshape(H, F)[:, :NFEAT] = w1
  th[OFF_W2:OFF_B2] = rng.normal(0.0, np.sqrt(2.0 / H), H * H)
  th[OFF_W3:OFF_B3] = rng.normal(0.0, np.sqrt(1.0 / H), H)
  return th


def forward(xbits: np.ndarray, wbits: np.ndarray):
  """x (S, 16) and weights as bf16 bit patterns -> (z, cache)."""
  x = from_bits(xbits)
  W1, b1, W2, b2, w3, b3 = unpack(from_bits(wbits))
  h1 = x @ W1.T + b1
  a1 = bf16_round(np.maximum(h1, 0.0))
  h2 = a1 @ W2.T + b2
  a2 = np.maximum(h2, 0.0)
  z = a2 @ w3 + b3
  return z, (x, h1, a1, h2, a2, W2, w3)


def loss_and_grad(xbits: np.ndarray, y: np.ndarray, wbits: np.ndarray):
  """Mean BCE-with-logits and its gradient (float64, device rounding points)."""
  z, (x, h1, a1, h2, a2, W2, w3) = forward(xbits, wbits)
  S = z.shape[0]
  y = np.asarray(y, dtype=np.float64)
  loss = float(np.mean(np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z)))))
  d3 = (1.0 / (1.0 + np.exp(-z)) - y) / S
  g = np.zeros(NPARAM, dtype=np.float64)
  g[OFF_W3:OFF_B3] = a2.T @ d3
  g[OFF_B3] = d3.sum()
  d2 = np.where(h2 > 0.0, d3[:, None] * w3[None, :], 0.0)
  d2b = bf16_round(d2)
  g[OFF_B2:OFF_W3] = d2b.sum(axis=0)
  g[OFF_W2:OFF_B2] = (d2b.T @ a1).reshape(-1)
  dh1 = d2b @ W2
  d1b = bf16_round(np.where(h1 > 0.0, dh1, 0.0))
  g[OFF_B1:OFF_W2] = d1b.sum(axis=0)
  gw1 = d1b.T @ x
  gw1[:, NFEAT:] = 0.0
  g[OFF_W1:OFF_B1] = gw1.reshape(-1)
  return loss, g


def float64_loss(x: np.ndarray, y: np.ndarray, theta: np.ndarray) -> float:
  """Pure float64 loss (no rounding) for finite-difference checks."""
  W1, b1, W2, b2, w3, b3 = unpack(theta)
  a1 = np.maximum(x @ W1.T + b1, 0.0)
  a2 = np.maximum(a1 @ W2.T + b2, 0.0)
  z = a2 @ w3 + b3
  return float(np.mean(np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z)))))


def float64_grad(x: np.ndarray, y: np.ndarray, theta: np.ndarray) -> np.ndarray:
  """Pure float64 analytic gradient (no rounding)."""
  W1, b1, W2, b2, w3, b3 = unpack(theta)
  h1 = x @ W1.T + b1
  a1 = np.maximum(h1, 0.0)
  h2 = a1 @ W2.T + b2
  a2 = np.maximum(h2, 0.0)
  z = a2 @ w3 + b3
  S = z.shape[0]
  d3 = (1.0 / (1.0 + np.exp(-z)) - y) / S
  g = np.zeros(NPARAM, dtype=np.float64)
  g[OFF_W3:OFF_B3] = a2.T @ d3
  g[OFF_B3] = d3.sum()
  d2 = np.where(h2 > 0.0, d3[:, None] * w3[None, :], 0.0)
  g[OFF_B2:OFF_W3] = d2.sum(axis=0)
  g[OFF_W2:OFF_B2] = (d2.T @ a1).reshape(-1)
  d1 = np.where(h1 > 0.0, d2 @ W2, 0.0)
  g[OFF_B1:OFF_W2] = d1.sum(axis=0)
  gw1 = d1.T @ x
  gw1[:, NFEAT:] = 0.0
  g[OFF_W1:OFF_B1] = gw1.reshape(-1)
  return g


def adam(theta, m, v, g, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
  """One Adam step in float32 arithmetic (Kingma & Ba 2015), the device's update order."""
  th = np.asarray(theta, np.float32).copy()
  g = np.asarray(g, np.float32).copy()
  pad = np.zeros(NPARAM, dtype=bool)
  pad[OFF_W1:OFF_B1] = (np.arange(H * F) % F) >= NFEAT
  g[pad] = 0.0
  if wd:
    wmask = np.zeros(NPARAM, dtype=bool)
    wmask[OFF_W1:OFF_B1] = True
    wmask[OFF_W2:OFF_B2] = True
    wmask[OFF_W3:OFF_B3] = True
    g = np.where(wmask, g + np.float32(wd) * th, g).astype(np.float32)
  one, b1f, b2f = np.float32(1), np.float32(b1), np.float32(b2)
  m = (b1f * m + (one - b1f) * g).astype(np.float32)
  v = (b2f * v + (one - b2f) * g * g).astype(np.float32)
  mh = m / (one - b1f ** np.float32(step))
  vh = v / (one - b2f ** np.float32(step))
  th = (th - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))).astype(np.float32)
  th[pad] = 0.0
  return th, m, v


def features(xy: np.ndarray, nn32: np.ndarray, D: np.ndarray) -> np.ndarray:
  """(n*32, 16) float32 features of (city, r-th neighbour), in the device's float32 order."""
  n = xy.shape[0]
  u = np.repeat(np.arange(n), 32)
  r = np.tile(np.arange(32), n)
  v = nn32.reshape(-1)
  f32 = np.float32
  d = D[u, v].astype(f32)
  du1 = np.maximum(D[u, nn32[u, 0]].astype(f32), f32(1))
  du8 = np.maximum(D[u, nn32[u, 7]].astype(f32), f32(1))
  du32 = np.maximum(D[u, nn32[u, 31]].astype(f32), f32(1))
  dv1 = np.maximum(D[v, nn32[v, 0]].astype(f32), f32(1))
  dv32 = np.maximum(D[v, nn32[v, 31]].astype(f32), f32(1))
  rev = np.full(n * 32, 32)
  hit = nn32[v] == u[:, None]
  has = hit.any(axis=1)
  rev[has] = hit[has].argmax(axis=1)
  out = np.zeros((n * 32, 16), dtype=f32)
  out[:, 0] = r.astype(f32) / f32(31)
  out[:, 1] = d / du32
  out[:, 2] = d / dv32
  out[:, 3] = du1 / np.maximum(d, f32(1))
  out[:, 4] = dv1 / np.maximum(d, f32(1))
  out[:, 5] = np.where(rev < 32, rev.astype(f32) / f32(31), f32(1.5))
  out[:, 6] = (rev < 32).astype(f32)
  out[:, 7] = (xy[v, 0] - xy[u, 0]).astype(f32) / du32
  out[:, 8] = (xy[v, 1] - xy[u, 1]).astype(f32) / du32
  out[:, 9] = d / du8
  out[:, 10] = (d - du1) / np.maximum(du32 - du1, f32(1))
  out[:, 11] = du8 / du32
  return out
