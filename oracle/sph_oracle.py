"""CPU oracle of the fused SPH density step (test infrastructure only).

Only tests/ and bench.py's checker legs may import this module.  The step
has no counterpart in the reference library: it is what the reference's
callers compute next from the k = 32 neighbours that knn_search
(cosmokdtree.f90:504-540) returns, a smoothing length and an SPH density
(the cosmology use named in README.md:18).  The GPU kernel
(cosmokdtree_amd/csrc/ckd_traverse.h, sph_row) computes it inside the kNN
traversal; this module computes it from the oracle's kNN rows
(oracle.OracleTree.knn), so the neighbour search itself is pinned by the
kNN parity of those rows and only the kernel sum is restated here.

* h = r_k / 2 (the cubic spline's support 2h reaches the k-th neighbour);
* rho = sum_j m_j W(r_j, h), W(r, h) = w(r / h) / (pi h^3),
  w(u) = 1 - 1.5 u^2 + 0.75 u^3 (0 <= u < 1), 0.25 (2 - u)^3 (1 <= u < 2),
  0 beyond (the M4 cubic spline, Monaghan & Lattanzio 1985);
* float64 arithmetic on the (float32 or float64) kNN distances.

Pinned by tests/test_sph_oracle.py against what the mathematics fixes: the
normalisation of W over space (integral of 4 pi u^2 w(u) du / pi = 1), the
continuity of w and w' at u = 1 and u = 2, the closed-form values at
u = 0, 1/2, 1, 3/2, and a brute-force density on a small point set.
"""
from __future__ import annotations

import numpy as np


def w_cubic(u):
    """M4 cubic spline shape w(u), u = r / h >= 0 (support u < 2)."""
    u = np.asarray(u, dtype=np.float64)
    inner = 1.0 - 1.5 * u * u + 0.75 * u * u * u
    outer = 0.25 * (2.0 - u) ** 3
    return np.where(u < 1.0, inner, np.where(u < 2.0, outer, 0.0))


def kernel(r, h):
    """W(r, h) = w(r / h) / (pi h^3): unit integral over 3-D space."""
    h = np.asarray(h, dtype=np.float64)
    return w_cubic(np.asarray(r, dtype=np.float64) / h) / (np.pi * h ** 3)


def sph_from_knn(dist, idx, mass=None):
    """(h, rho) of each query from its kNN rows dist (nq, k), idx (nq, k)
    (sorted or not; every slot filled).  mass: per point or None (unit)."""
    d = np.asarray(dist, dtype=np.float64)
    h = d.max(axis=1) * 0.5
    m = 1.0 if mass is None else np.asarray(mass, dtype=np.float64)[np.asarray(idx)]
    rho = (m * kernel(d, h[:, None])).sum(axis=1)
    return h, rho


def sph_brute(points, queries, k, mass=None, boxsize=None):
    """(h, rho) by brute force: all pairwise distances (minimum image when
    periodic), the k smallest per query.  For small inputs only."""
    p = np.asarray(points, dtype=np.float64)
    q = np.asarray(queries, dtype=np.float64)
    diff = q[:, None, :] - p[None, :, :]
    if boxsize is not None:
        L = np.asarray(boxsize, dtype=np.float64)
        a = np.abs(diff)
        diff = np.minimum(a, L - a)
    r = np.sqrt((diff * diff).sum(axis=2))
    nn = np.argsort(r, axis=1, kind="stable")[:, :k]
    rk = np.take_along_axis(r, nn, axis=1)
    return sph_from_knn(rk, nn, mass)
