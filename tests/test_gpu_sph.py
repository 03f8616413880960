"""The fused SPH density step (ckd_sph_density) against the oracle.

h must equal half of the oracle's k-th kNN distance bit for bit (the kernel
computes it from the same bit-exact distance); rho is a float sum of k
positive terms in a different order from the oracle's float64 sum, so it is
compared with a relative tolerance of 1e-5 in float32 and 1e-12 in float64
(k <= 64 terms, each a few roundings).  Points tied with the k-th neighbour
lie at u = 2 where the kernel vanishes, so the choice among them does not
change rho.
"""
import numpy as np
import pytest
import torch

from oracle import sph_oracle as so
from oracle.oracle import OracleTree

pytestmark = pytest.mark.gpu

RTOL = {np.float32: 1e-5, np.float64: 1e-12}


def _clustered(n, seed, dtype):
    from cosmokdtree_amd.synthetic import clustered_periodic
    return clustered_periodic(n, 100.0, 0.5, seed=seed, dtype=dtype), np.full(3, 100.0)


CASES = {
    "clu_f32_leaf16": lambda: (*_clustered(60000, 21, np.float32), 16),
    "clu_f64_leaf8": lambda: (*_clustered(30000, 22, np.float64), 8),
    "uni_f32_leaf32": lambda: (np.random.default_rng(23).random((50000, 3)).astype(np.float32), None, 32),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def scase(request, ckd):
    pts, L, leaf = CASES[request.param]()
    print(f"\n[sph {request.param}] n={len(pts)} leaf={leaf}", flush=True)
    tree = ckd.kdtree._build(pts, leaf, L)
    orc = OracleTree(pts, leaf, L)
    rng = np.random.default_rng(24)
    # half the queries are tree points (they count themselves), half moved
    q = pts[rng.integers(0, len(pts), 6000)].astype(np.float64)
    q[3000:] += (rng.random((3000, 3)) - 0.5) * (0.3 if L is not None else 0.003)
    if L is not None:
        q = np.clip(q, -L / 2, L / 2)
    q = q.astype(pts.dtype)
    mass = (rng.random(len(pts)) + 0.5).astype(pts.dtype)
    yield dict(pts=pts, L=L, tree=tree, orc=orc, q=q, mass=mass)
    tree.free()
    orc.close()


@pytest.mark.parametrize("k", [8, 32, 64])
@pytest.mark.parametrize("masses", [False, True])
@pytest.mark.parametrize("nq", [6000, 700])  # leaf-ordered and input-ordered launches
def test_sph_matches_oracle(scase, k, masses, nq):
    q = scase["q"][:nq]
    mass = scase["mass"] if masses else None
    h, rho = scase["tree"].sph_density(q, k, mass)
    h, rho = h.cpu().numpy(), rho.cpu().numpy()
    d_ref, i_ref = scase["orc"].knn(q, k, True)
    dt = scase["pts"].dtype.type
    np.testing.assert_array_equal(h, (d_ref[:, -1] * dt(0.5)).astype(dt))
    h_o, rho_o = so.sph_from_knn(d_ref, i_ref, mass)
    np.testing.assert_allclose(rho, rho_o, rtol=RTOL[dt], atol=0)
    # the GPU's own kNN rows give the same h
    d_gpu, _ = scase["tree"].knn(q, k, True)
    np.testing.assert_array_equal(h, d_gpu[:, -1].cpu().numpy() * dt(0.5))


def test_sph_brute_force(ckd):
    rng = np.random.default_rng(25)
    pts = (rng.random((3000, 3)) - 0.5) * 10.0
    L = np.full(3, 10.0)
    q = np.clip(pts[:200] + (rng.random((200, 3)) - 0.5) * 0.2, -5, 5)
    mass = rng.random(3000) + 0.5
    tree = ckd.kdtree._build(pts, 8, L)
    try:
        h, rho = tree.sph_density(q, 32, mass)
        h_b, rho_b = so.sph_brute(pts, q, 32, mass, L)
        np.testing.assert_allclose(h.cpu().numpy(), h_b, rtol=1e-14)
        np.testing.assert_allclose(rho.cpu().numpy(), rho_b, rtol=1e-12)
    finally:
        tree.free()


def test_sph_rejects(ckd):
    rng = np.random.default_rng(26)
    t2 = ckd.kdtree._build(rng.random((500, 2)).astype(np.float32), 8, None)
    t3 = ckd.kdtree._build(rng.random((500, 3)).astype(np.float32), 8, None)
    try:
        with pytest.raises(ValueError):
            t2.sph_density(rng.random((10, 2)), 8)
        for k in (1, 65, 501):
            with pytest.raises(ValueError):
                t3.sph_density(rng.random((10, 3)), k)
        with pytest.raises(ValueError):
            t3.sph_density(rng.random((10, 3)), 8, np.ones(10))
        h, rho = t3.sph_density(np.zeros((0, 3)), 8)
        assert h.numel() == 0 and rho.numel() == 0
    finally:
        t2.free()
        t3.free()


@pytest.mark.parametrize("k", [8, 32, 64])
def test_sph_unit_masses_paths_agree(scase, k):
    """Without masses the walk keeps distances only (no rows); with explicit
    unit masses it keeps rows and looks the masses up.  The lists hold the
    same distances in the same order, so h and rho agree bit for bit."""
    q = scase["q"]
    h0, r0 = scase["tree"].sph_density(q, k)
    h1, r1 = scase["tree"].sph_density(q, k, np.ones(len(scase["pts"]), dtype=scase["pts"].dtype))
    assert torch.equal(h0, h1) and torch.equal(r0, r1)
