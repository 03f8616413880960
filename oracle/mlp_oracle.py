"""CPU oracle of the learned search-radius network (test infrastructure only).

Only tests/ and bench.py's checker legs may import this module.  It has no
reference counterpart (the reference has no learned component); it restates,
in numpy float64, exactly what the GPU kernels in
cosmokdtree_amd/csrc/ckd_mlp.hip compute, so their forward pass and
gradients can be checked:

* features: 30 per query = 3 x ((q + L/2) / L) followed by the 27
  log1p(counts) of the 3x3x3 periodic neighbourhood of the query's cell in a
  G^3 occupancy grid (G = 64), ordered dz-major (dz, dy, dx in -1..1), plus
  2 zero pads (32 columns);
* network: y = w3 . gelu(W2^T gelu(W1^T x + b1) + b2) + b3, widths
  32 -> 64 -> 64 -> 1, GELU(z) = z * 0.5 * (1 + erf(z / sqrt 2));
* precision: inputs, weights and hidden activations rounded to bfloat16
  (round to nearest even) before every product, products accumulated in
  float32 or wider, biases and the output in float32;
* loss: mean squared error against log(r_k); Adam (beta1 0.9, beta2 0.999,
  eps 1e-8) on float32 master weights.

The parameter vector is the flat float32 layout used by the GPU kernels:
W1[32,64] (row = input feature), b1[64], W2[64,64] (row = input unit),
b2[64], w3[64], b3[1].
"""
from __future__ import annotations

from math import sqrt

import numpy as np
from scipy.special import erf

NIN, NH = 32, 64
SIZES = dict(W1=NIN * NH, b1=NH, W2=NH * NH, b2=NH, w3=NH, b3=1)
NPARAM = sum(SIZES.values())  # 6337
G_DEFAULT = 64


def offsets():
    o, out = 0, {}
    for k, n in SIZES.items():
        out[k] = (o, o + n)
        o += n
    return out


OFF = offsets()


def unpack(p):
    p = np.asarray(p)
    return dict(W1=p[OFF["W1"][0]:OFF["W1"][1]].reshape(NIN, NH), b1=p[OFF["b1"][0]:OFF["b1"][1]],
                W2=p[OFF["W2"][0]:OFF["W2"][1]].reshape(NH, NH), b2=p[OFF["b2"][0]:OFF["b2"][1]],
                w3=p[OFF["w3"][0]:OFF["w3"][1]], b3=p[OFF["b3"][0]:OFF["b3"][1]])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return z * 0.5 * (1.0 + erf(z / sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / sqrt(2.0))) + z * np.exp(-0.5 * z * z) / sqrt(2.0 * np.pi)


def init_params(seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(NPARAM, dtype=np.float32)
    w = unpack(p)
    w["W1"][:30] = rng.standard_normal((30, NH)) * sqrt(2.0 / 30)
    w["W2"][:] = rng.standard_normal((NH, NH)) * sqrt(2.0 / NH)
    w["w3"][:] = rng.standard_normal(NH) * sqrt(1.0 / NH)
    return p


def occupancy(points, L, G=G_DEFAULT):
    """Counts of points per cell of a G^3 grid over [-L/2, L/2)^3."""
    c = np.floor((np.asarray(points, np.float64) + L / 2) / L * G).astype(np.int64)
    c = np.clip(c, 0, G - 1)
    flat = (c[:, 2] * G + c[:, 1]) * G + c[:, 0]
    return np.bincount(flat, minlength=G ** 3).astype(np.int32)


def features(q, L, grid, G=G_DEFAULT):
    q = np.asarray(q, np.float64)
    f = np.zeros((len(q), NIN), dtype=np.float32)
    f[:, :3] = ((q + L / 2) / L).astype(np.float32)
    c = np.clip(np.floor((q + L / 2) / L * G).astype(np.int64), 0, G - 1)
    j = 3
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                x = (c[:, 0] + dx) % G
                y = (c[:, 1] + dy) % G
                z = (c[:, 2] + dz) % G
                f[:, j] = np.log1p(grid[(z * G + y) * G + x].astype(np.float64)).astype(np.float32)
                j += 1
    return f


def forward(p, X, rounding=True):
    """Returns y (float64) and the intermediates for backward.
    rounding=False skips every bfloat16 rounding (exact float64 network),
    which lets finite differences check the backward formulas."""
    R = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    w = unpack(np.asarray(p, np.float32 if rounding else np.float64))
    Xb = R(X)
    z1 = Xb @ R(w["W1"]) + w["b1"].astype(np.float64)
    h1 = R(gelu(z1))
    z2 = h1 @ R(w["W2"]) + w["b2"].astype(np.float64)
    h2 = R(gelu(z2))
    y = h2 @ R(w["w3"]) + np.float64(w["b3"][0])
    return y, dict(Xb=Xb, z1=z1, h1=h1, z2=z2, h2=h2, w=w, R=R)


def loss_and_grad(p, X, t, rounding=True):
    """MSE loss and its gradient w.r.t. the float32 parameter vector.
    Backward products use the same bfloat16-rounded operands the GPU uses
    (activations, weights and upstream gradients rounded before a product)."""
    y, c = forward(p, X, rounding)
    bf16 = c["R"]
    B = len(t)
    r = y - np.asarray(t, np.float64)
    loss = float(np.mean(r * r))
    dy = 2.0 * r / B
    w = c["w"]
    g = np.zeros(NPARAM, dtype=np.float64)
    gw = unpack(g)
    gw["b3"][:] = dy.sum()
    gw["w3"][:] = c["h2"].T @ dy
    dz2 = np.outer(dy, bf16(w["w3"])) * gelu_grad(c["z2"])
    gw["b2"][:] = dz2.sum(axis=0)
    dz2b = bf16(dz2)
    gw["W2"][:] = c["h1"].T @ dz2b
    dz1 = (dz2b @ bf16(w["W2"]).T) * gelu_grad(c["z1"])
    gw["b1"][:] = dz1.sum(axis=0)
    gw["W1"][:] = c["Xb"].T @ bf16(dz1)
    return loss, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32 (bias-corrected), as the GPU kernel does."""
    p = np.asarray(p, np.float32).copy()
    g = np.asarray(g, np.float32)
    m = (np.float32(b1) * m + np.float32(1 - b1) * g).astype(np.float32)
    v = (np.float32(b2) * v + np.float32(1 - b2) * g * g).astype(np.float32)
    mh = m / np.float32(1 - b1 ** step)
    vh = v / np.float32(1 - b2 ** step)
    p -= np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))
    return p, m, v
