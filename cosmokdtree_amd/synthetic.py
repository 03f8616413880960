"""Seeded synthetic point sets of the benchmark workloads.

* ``uniform(n, d, seed)``: i.i.d. uniform in [0, 1)^d.
* ``clustered_periodic(n, L, halo_frac, seed)``: the clustered periodic box.
  A fraction ``1 - halo_frac`` of the points is uniform in [-L/2, L/2)^3;
  the rest sits in halos with uniformly placed centres and sizes drawn from
  dN/dn ~ n^-1.9 on [10, 1e5] (inverse CDF, drawn until the budget is
  filled, the last halo truncated), each an isotropic Gaussian of
  sigma = 0.02 n^(1/3) wrapped into the box.  Coordinates lie in
  [-L/2, L/2] as a periodic tree requires (cosmokdtree.f90:176-190).
* ``halo_set(n, L, seed)``: one Gaussian halo plus a 20% uniform background
  (the batched uneven-instance workload).
* ``batched_sets(total, nsets, L, seed)``: the batched workload at scale:
  nsets set sizes log-uniform in [1e3, 1e6] rescaled to sum to `total`,
  each set a halo_set-style Gaussian halo (sigma = 0.02 n_halo^(1/3)) plus
  20% uniform background in its own periodic box of side L; 10% of each
  set's points as its queries, k per set from {8, 16, 32, 64}, a ball
  radius per set log-uniform in [0.05, 0.5].

All generators use numpy's PCG64 with the given seed; float64 results can
be cast to float32 by the caller.
"""
from __future__ import annotations

import numpy as np

SEED_POINTS = 20261203
SEED_QUERIES = SEED_POINTS + 1
SEED_SUBSET = SEED_POINTS + 2
SEED_TRAIN = SEED_POINTS + 3


def uniform(n: int, d: int = 3, seed: int = SEED_POINTS) -> np.ndarray:
    return np.random.Generator(np.random.PCG64(seed)).random((n, d))


def wrap(x: np.ndarray, L: float) -> np.ndarray:
    return x - L * np.floor((x + L / 2) / L)


def halo_sizes(budget: int, rng: np.random.Generator, nmin: float = 10.0, nmax: float = 1e5,
               slope: float = 1.9) -> np.ndarray:
    """Halo occupations from dN/dn ~ n^-slope on [nmin, nmax], summing to budget."""
    a = 1.0 - slope  # -0.9
    lo, hi = nmin ** a, nmax ** a
    sizes = []
    total = 0
    while total < budget:
        u = rng.random(max(1024, (budget - total) // 50))
        n = np.floor((lo - u * (lo - hi)) ** (1.0 / a)).astype(np.int64)
        n = np.clip(n, int(nmin), int(nmax))
        cs = np.cumsum(n)
        take = np.searchsorted(cs, budget - total, side="left")
        if take < len(n):
            n = n[:take + 1].copy()
            n[-1] -= cs[take] - (budget - total)
            n = n[n > 0]
            sizes.append(n)
            total = budget
        else:
            sizes.append(n)
            total += int(cs[-1])
    return np.concatenate(sizes)


def clustered_periodic(n: int, L: float = 100.0, halo_frac: float = 0.5, seed: int = SEED_POINTS,
                       dtype=np.float64) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    n_halo = int(round(n * halo_frac))
    n_uni = n - n_halo
    out = np.empty((n, 3), dtype=np.float64)
    out[:n_uni] = (rng.random((n_uni, 3)) - 0.5) * L
    if n_halo:
        sizes = halo_sizes(n_halo, rng)
        centres = (rng.random((len(sizes), 3)) - 0.5) * L
        sig = 0.02 * np.cbrt(sizes.astype(np.float64))
        rep_c = np.repeat(centres, sizes, axis=0)
        rep_s = np.repeat(sig, sizes)[:, None]
        out[n_uni:] = wrap(rep_c + rep_s * rng.standard_normal((n_halo, 3)), L)
    out = out.astype(dtype)
    h = dtype(L / 2) if dtype == np.float32 else L / 2
    np.clip(out, -h, h, out=out)
    return out


def halo_set(n: int, L: float = 10.0, seed: int = 0, dtype=np.float32, bg_frac: float = 0.2) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = int(round(n * bg_frac))
    nh = n - nb
    c = (rng.random(3) - 0.5) * L
    sig = 0.02 * np.cbrt(nh)
    pts = np.concatenate([(rng.random((nb, 3)) - 0.5) * L, wrap(c + sig * rng.standard_normal((nh, 3)), L)])
    pts = pts.astype(dtype)
    h = dtype(L / 2)
    np.clip(pts, -h, h, out=pts)
    return pts


def batched_sets(total: int = 200_000_000, nsets: int = 4096, L: float = 10.0, seed: int = SEED_POINTS,
                 query_frac: float = 0.1):
    """The batched workload, generated vectorised in float32.  Returns
    (points (N, 3) float32, offsets (nsets+1) int64, queries (Q, 3) float32,
    query offsets (nsets+1) int64, k (nsets) int32, radius (nsets) float64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    raw = np.exp(rng.uniform(np.log(1e3), np.log(1e6), nsets))
    sizes = np.maximum(np.floor(raw * total / raw.sum()).astype(np.int64), 1)
    sizes[-1] += total - sizes.sum()
    off = np.zeros(nsets + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    nb = np.round(sizes * 0.2).astype(np.int64)
    nh = sizes - nb
    centres = ((rng.random((nsets, 3)) - 0.5) * L).astype(np.float32)
    sig = (0.02 * np.cbrt(np.maximum(nh, 1).astype(np.float64))).astype(np.float32)
    set_of = np.repeat(np.arange(nsets), sizes)
    local = np.arange(total, dtype=np.int64) - off[set_of]
    halo = local >= nb[set_of]
    pts = ((rng.random((total, 3), dtype=np.float32) - np.float32(0.5)) * np.float32(L))
    g = rng.standard_normal((total, 3), dtype=np.float32)
    hp = centres[set_of] + sig[set_of, None] * g
    hp -= np.float32(L) * np.floor((hp + np.float32(L / 2)) / np.float32(L))
    pts[halo] = hp[halo]
    del g, hp
    h = np.float32(L / 2)
    np.clip(pts, -h, h, out=pts)
    nq = np.maximum((sizes * query_frac).astype(np.int64), 1)
    qoff = np.zeros(nsets + 1, dtype=np.int64)
    np.cumsum(nq, out=qoff[1:])
    qsel = off[np.repeat(np.arange(nsets), nq)] + (rng.random(int(qoff[-1])) * np.repeat(sizes, nq)).astype(np.int64)
    queries = pts[qsel]
    k = rng.choice(np.array([8, 16, 32, 64], dtype=np.int32), nsets)
    radius = np.exp(rng.uniform(np.log(0.05), np.log(0.5), nsets))
    return pts, off, queries, qoff, k, radius
