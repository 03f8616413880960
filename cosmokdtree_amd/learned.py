"""Learned search-radius proposal for kNN on periodic 3-D trees.

A small network predicts, per query, the log of its k-th neighbour
distance from where the query sits and how crowded its surroundings are:
32 features = the query position scaled to [0, 1) per axis, log1p of the
point counts in the 27 cells around it in a G^3 occupancy grid (G = 64),
and two local densities read off the tree itself (the tight point boxes of
the smallest ancestors of the query's leaf holding 16 and 64 points).
The network (32 -> 64 -> 64 -> 1, GELU) runs on bf16 MFMA with fp32
accumulation, and is trained on the GPU by the library's own forward,
backward and Adam kernels (include/cosmokdtree_amd.h, ckd_mlp_*), on
(query, true k-th neighbour distance) pairs that a tree built on an
independent point set produces.

``margin * exp(prediction)`` seeds the kNN pruning bound (ckd_knn's
seed_r).  A query that finds fewer than k points inside its seed is searched
again without it, so results are those of the plain search whatever the
network says; the margin trades pruning (small seeds) against re-runs
(seeds that are too small).

No reference counterpart: the reference has no learned component.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from ._lib import check, lib
from .kdtree import KDTree, _stream, _vp

lib.ckd_mlp_nparam.restype = C.c_int
lib.ckd_mlp_nparam.argtypes = []
lib.ckd_occupancy.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
lib.ckd_mlp_table.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
lib.ckd_mlp_features.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p]
lib.ckd_knn_learned.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p, C.c_float,
                                C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
lib.ckd_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_int32, C.c_void_p, C.c_void_p]
lib.ckd_mlp_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
lib.ckd_mlp_adam.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_void_p]

NPARAM = lib.ckd_mlp_nparam()
NIN, NH = 32, 64
GRID = 64


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation of the flat fp32 parameter vector."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(NPARAM, dtype=np.float32)
    o = 0
    W1 = p[o:o + NIN * NH].reshape(NIN, NH)
    W1[:] = rng.standard_normal((NIN, NH)) * math.sqrt(2.0 / NIN)
    o += NIN * NH + NH
    p[o:o + NH * NH] = (rng.standard_normal((NH, NH)) * math.sqrt(2.0 / NH)).ravel()
    o += NH * NH + NH
    p[o:o + NH] = rng.standard_normal(NH) * math.sqrt(1.0 / NH)
    return p


def occupancy(tree: KDTree, G: int = GRID) -> torch.Tensor:
    grid = torch.empty(G ** 3, dtype=torch.int32, device=tree.device)
    check(lib.ckd_occupancy(tree._ptr, G, _vp(grid), _stream()), "occupancy")
    return grid


def feature_table(grid: torch.Tensor, G: int = GRID) -> torch.Tensor:
    """bf16 log1p of every grid count (int16 bit patterns)."""
    tab = torch.empty(G ** 3, dtype=torch.int16, device=grid.device)
    check(lib.ckd_mlp_table(_vp(grid), G, _vp(tab), _stream()), "feature table")
    return tab


def features(tree: KDTree, queries: torch.Tensor, table: torch.Tensor, G: int = GRID, leaf_of=None) -> torch.Tensor:
    """(nq, 32) bf16 features (int16 bit patterns): 3 positions, 27 grid
    neighbourhood counts, 2 tree density features."""
    q = tree._queries(queries)
    X = torch.empty((q.shape[0], NIN), dtype=torch.int16, device=q.device)
    check(lib.ckd_mlp_features(tree._ptr, _vp(q), q.shape[0], G, _vp(table), _vp(leaf_of), _vp(X), _stream()),
          "features")
    return X


class RadiusModel:
    """The network's fp32 master parameters and Adam state on the GPU."""

    def __init__(self, params=None, seed: int = 0, device="cuda"):
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        self.params = torch.from_numpy(p.copy()).to(device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.loss = torch.zeros(1, dtype=torch.float32, device=device)
        self.step = 0
        self.margin = 1.0

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        """Predicted log radius per row."""
        out = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        check(lib.ckd_mlp_forward(_vp(self.params), _vp(X), X.shape[0], 1.0, 0, _vp(out), _stream()), "forward")
        return out

    def seeds(self, X: torch.Tensor, margin: float | None = None) -> torch.Tensor:
        """Seed radii margin * exp(prediction), fused in the forward kernel."""
        out = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        m = self.margin if margin is None else margin
        check(lib.ckd_mlp_forward(_vp(self.params), _vp(X), X.shape[0], float(m), 1, _vp(out), _stream()), "seeds")
        return out

    def grad_and_loss(self, X: torch.Tensor, t: torch.Tensor):
        """MSE loss (mean) and gradient for one batch (the GPU kernels)."""
        check(lib.ckd_mlp_grad(_vp(self.params), _vp(X), _vp(t), X.shape[0], _vp(self.grad), _vp(self.loss),
                               _stream()), "grad")
        return self.loss / X.shape[0], self.grad

    def adam(self, lr: float):
        self.step += 1
        check(lib.ckd_mlp_adam(_vp(self.params), _vp(self.grad), _vp(self.m), _vp(self.v), self.step, float(lr),
                               _stream()), "adam")

    def train(self, X: torch.Tensor, t: torch.Tensor, epochs: int = 4, batch: int = 4096, lr: float = 2e-3,
              seed: int = 0, log=None):
        """Adam on shuffled mini-batches; returns per-epoch mean losses."""
        n = X.shape[0]
        g = torch.Generator(device=X.device)
        g.manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, device=X.device, generator=g)
            Xs, ts = X[perm], t[perm]
            acc = torch.zeros(1, dtype=torch.float32, device=X.device)
            nb = 0
            lr_ep = lr * (0.5 ** ep)
            for b0 in range(0, n - batch + 1, batch):
                loss, _ = self.grad_and_loss(Xs[b0:b0 + batch], ts[b0:b0 + batch])
                self.adam(lr_ep)
                acc += loss
                nb += 1
            losses.append(float(acc.item() / max(nb, 1)))
            if log:
                log(f"epoch {ep}: mse {losses[-1]:.5f}")
        return losses

    def calibrate(self, X: torch.Tensor, r_true: torch.Tensor, quantile: float = 0.98) -> float:
        """Margin such that `quantile` of the rows have r_true <= margin*exp(pred)."""
        ratio = torch.log(r_true) - self.forward(X)
        self.margin = float(torch.exp(torch.quantile(ratio.float(), quantile)).item())
        return self.margin

    def save(self, path: str):
        np.savez(path, params=self.params.cpu().numpy(), margin=np.array(self.margin))

    @classmethod
    def load(cls, path: str, device="cuda"):
        z = np.load(path)
        m = cls(params=z["params"], device=device)
        m.margin = float(z["margin"])
        return m


def training_pairs(tree: KDTree, queries: torch.Tensor, k: int, G: int = GRID):
    """(features, log r_k, r_k) for queries against `tree` (exact kNN)."""
    d, _ = tree.knn(queries, k, sorted=True)
    r = d[:, k - 1].float()
    X = features(tree, queries, feature_table(occupancy(tree, G), G), G)
    return X, torch.log(r), r


def fit(train_tree: KDTree, n_train: int, k: int = 32, epochs: int = 4, batch: int = 4096, lr: float = 2e-3,
        quantile: float = 0.98, seed: int = 0, log=None) -> RadiusModel:
    """Train a model on queries drawn from the tree's own points and
    calibrate its margin on a held-out tenth."""
    t0 = time.time()
    g = torch.Generator(device=train_tree.device)
    g.manual_seed(seed + 1)
    pts = train_tree_points(train_tree)
    sel = torch.randint(0, pts.shape[0], (n_train,), device=pts.device, generator=g)
    q = pts[sel].contiguous()
    X, logr, r = training_pairs(train_tree, q, k)
    n_fit = n_train - n_train // 10
    model = RadiusModel(seed=seed)
    model.train(X[:n_fit], logr[:n_fit], epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    model.calibrate(X[n_fit:], r[n_fit:], quantile)
    if log:
        log(f"trained on {n_fit} pairs in {time.time() - t0:.1f}s, margin {model.margin:.3f}")
    return model


def train_tree_points(tree: KDTree) -> torch.Tensor:
    if getattr(tree, "points", None) is None:
        raise ValueError("the training tree must keep its points (build with kdtree._build(..., keep_points=True))")
    return tree.points


def seeded_knn(tree: KDTree, model: RadiusModel, queries: torch.Tensor, k: int, sorted: bool = True,
               table: torch.Tensor | None = None, stats=None, margin: float | None = None):
    """kNN with learned seeds in one library call (ckd_knn_learned): leaves,
    features, network, seeded traversal and the re-search of short seeds.
    `table` is feature_table(occupancy(tree)), computed once per tree."""
    tree._check_alive()
    if table is None:
        table = feature_table(occupancy(tree))
    q = tree._queries(queries)
    nq = q.shape[0]
    dist = torch.empty((nq, k), dtype=tree.torch_dtype, device=tree.device)
    idx = torch.empty((nq, k), dtype=torch.int64, device=tree.device)
    if nq == 0:
        return dist, idx
    m = model.margin if margin is None else margin
    check(lib.ckd_knn_learned(tree._ptr, _vp(q), nq, int(k), int(bool(sorted)), _vp(model.params), float(m), GRID,
                              _vp(table), _vp(dist), _vp(idx), _vp(stats), _stream()), "knn_learned")
    return dist, idx
