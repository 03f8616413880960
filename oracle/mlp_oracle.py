"""CPU oracle of the learned search-radius network (test infrastructure only).

Only tests/ and bench.py's checker legs may import this module.  It has no
reference counterpart (the reference has no learned component); it restates,
in numpy float64, exactly what the GPU kernels in
cosmokdtree_amd/csrc/ckd_mlp.hip compute, so their forward pass and
gradients can be checked:

* features: 32 per query = 3 x ((q + L/2) / L), the 27 log1p(counts) of
  the 3x3x3 periodic neighbourhood of the query's cell in a G^3 occupancy
  grid (G = 64), ordered dz-major (dz, dy, dx in -1..1), and 2 tree density
  features (tree_features: (1/3) ln(count / volume) of the tight point box
  of the smallest ancestor of the query's leaf holding >= 16, then >= 64
  points; the leaf is found by the first descent of the kNN search,
  cosmokdtree.f90:820-836, on the oracle's own tree);
* network: y = w3 . gelu(W2^T gelu(W1^T x + b1) + b2) + b3, widths
  32 -> 64 -> 64 -> 1, GELU in its sigmoid form z * sigmoid(1.702 z);
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
from scipy.special import expit

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


GELU_A = 1.702


def gelu(z):
    """Sigmoid form of GELU, z * sigmoid(1.702 z) (Hendrycks & Gimpel 2016)."""
    return z * expit(GELU_A * z)


def gelu_grad(z):
    s = expit(GELU_A * z)
    return s + GELU_A * z * s * (1.0 - s)


def init_params(seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(NPARAM, dtype=np.float32)
    w = unpack(p)
    w["W1"][:] = rng.standard_normal((NIN, NH)) * sqrt(2.0 / NIN)
    w["W2"][:] = rng.standard_normal((NH, NH)) * sqrt(2.0 / NH)
    w["w3"][:] = rng.standard_normal(NH) * sqrt(1.0 / NH)
    return p


def occupancy(points, L, G=G_DEFAULT):
    """Counts of points per cell of a G^3 grid over [-L/2, L/2)^3."""
    c = np.floor((np.asarray(points, np.float64) + L / 2) / L * G).astype(np.int64)
    c = np.clip(c, 0, G - 1)
    flat = (c[:, 2] * G + c[:, 1]) * G + c[:, 0]
    return np.bincount(flat, minlength=G ** 3).astype(np.int32)


def tree_features(dump, q, L, thresholds=(16, 64)):
    """Columns 30-31: density features from an oracle tree dump
    (oracle.OracleTree.dump: preorder nodes, each with a contiguous range
    [start, start + count) of the leaf-ordered points lpts).  `q` must be in
    the tree's precision (the descent compares q[axis] < split in it)."""
    q = np.asarray(q)
    L = np.broadcast_to(np.asarray(L, np.float64), (3,))
    leaf = dump["is_leaf"].astype(bool)
    axis, split = dump["axis"], dump["split"]
    left, right = dump["left"], dump["right"]
    start, count, lpts = dump["start"], dump["count"], dump["lpts"]
    parent = np.full(len(axis), -1, np.int64)
    inner = np.nonzero(~leaf)[0]
    parent[left[inner]] = inner
    parent[right[inner]] = inner
    v = np.zeros(len(q), np.int64)
    rows = np.arange(len(q))
    while True:  # first descent: left when q[axis] < split
        act = ~leaf[v]
        if not act.any():
            break
        va = v[act]
        goleft = q[rows[act], axis[va]] < split[va]
        v[act] = np.where(goleft, left[va], right[va])
    out = np.zeros((len(q), 2), np.float32)
    for col, c in enumerate(thresholds):
        while True:
            up = (count[v] < c) & (parent[v] >= 0)
            if not up.any():
                break
            v[up] = parent[v[up]]
        for node in np.unique(v):
            sel = v == node
            pts = lpts[start[node]:start[node] + count[node]].astype(np.float64)
            ext = np.maximum(pts.max(axis=0) - pts.min(axis=0), 1e-6 * L)
            V = (ext[0] * ext[1]) * ext[2]
            out[sel, col] = np.float32(np.log(count[node] / V) / 3.0)
    return out


def features(q, L, grid, G=G_DEFAULT, tree_cols=None):
    """The 32 feature columns; tree_cols = tree_features(...) fills 30-31
    (zero otherwise)."""
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
    if tree_cols is not None:
        f[:, 30:32] = tree_cols
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
