"""Pins of the SPH density oracle (oracle/sph_oracle.py) by what the
mathematics fixes: normalisation, continuity, closed-form values, brute
force on a lattice.  CPU only."""
import numpy as np
from scipy.integrate import quad

from oracle import sph_oracle as so


def test_kernel_normalised():
    # integral over space of W(r, h) = 4 pi int_0^2h r^2 W dr = 1 for any h
    for h in (0.1, 1.0, 3.7):
        val = 4 * np.pi * (quad(lambda r: r * r * so.kernel(r, h), 0, h)[0]
                           + quad(lambda r: r * r * so.kernel(r, h), h, 2 * h)[0])
        assert abs(val - 1.0) < 1e-10


def test_shape_closed_forms_and_continuity():
    assert so.w_cubic(0.0) == 1.0
    assert so.w_cubic(0.5) == 1.0 - 1.5 / 4 + 0.75 / 8
    assert so.w_cubic(1.0) == 0.25
    assert so.w_cubic(1.5) == 0.25 * 0.125
    assert so.w_cubic(2.0) == 0.0 and so.w_cubic(5.0) == 0.0
    e = 1e-7
    for u in (1.0, 2.0):  # w and w' continuous at the joins
        assert abs(so.w_cubic(u - e) - so.w_cubic(u + e)) < 1e-6
        dl = (so.w_cubic(u - e) - so.w_cubic(u - 2 * e)) / e
        dr = (so.w_cubic(u + 2 * e) - so.w_cubic(u + e)) / e
        assert abs(dl - dr) < 1e-5
    # w'(u) = -3u + 2.25u^2 below 1: slope -0.75 at u = 1
    assert abs((so.w_cubic(1.0) - so.w_cubic(1.0 - e)) / e + 0.75) < 1e-5


def test_brute_force_on_lattice():
    # a periodic cubic lattice of spacing a, query at a lattice point: the
    # 7 nearest are the point and its 6 face neighbours (r = a) and the
    # 8th..19th are the 12 edge neighbours at sqrt(2) a
    a, m = 1.0, 8
    g = np.arange(m) * a - m * a / 2
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    L = np.full(3, m * a)
    h, rho = so.sph_brute(pts, pts[:1], 7, boxsize=L)
    assert h[0] == 0.5 * a
    # r = a = 2h for the 6 faces: only the point itself contributes w(0) = 1
    assert abs(rho[0] - 1.0 / (np.pi * h[0] ** 3)) < 1e-12
    h, rho = so.sph_brute(pts, pts[:1], 19, boxsize=L)
    hh = np.sqrt(2) * a / 2
    want = (1.0 + 6 * so.w_cubic(a / hh)) / (np.pi * hh ** 3)
    assert abs(h[0] - hh) < 1e-15 and abs(rho[0] - want) < 1e-12 * want


def test_from_knn_matches_brute_and_masses():
    rng = np.random.default_rng(5)
    pts = rng.random((400, 3))
    q = rng.random((50, 3))
    mass = rng.random(400) + 0.5
    h, rho = so.sph_brute(pts, q, 16, mass)
    r = np.sqrt(((q[:, None] - pts[None]) ** 2).sum(-1))
    nn = np.argsort(r, axis=1)[:, :16]
    # hand sum, one query at a time
    for j in range(5):
        hj = r[j, nn[j, -1]] / 2
        s = sum(mass[i] * so.w_cubic(r[j, i] / hj) for i in nn[j]) / (np.pi * hj ** 3)
        assert abs(h[j] - hj) < 1e-15 and abs(rho[j] - s) < 1e-12 * s
