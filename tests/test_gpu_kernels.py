"""Every kNN kernel (per query, per wave with parent pointers, per wave
over wide records with an LDS stack) against the oracle.

The queries are numerous enough (>= 4096) to run in leaf order, as large
batches do.  Cases cover leaves longer than the per-lane candidate buffer
(leaf 40), both stack depths, a tree deeper than either stack (the stack
kernel must hand over to the parent-pointer walk), seeded searches with
short seeds (retry list), and the query counter under retries.
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree
from tests.parity import assert_knn_rows

pytestmark = pytest.mark.gpu

KERNELS = [("lane", 0, 0), ("wave", 0, 0), ("stack", 32, 0), ("stack", 64, 0), ("stack", 0, 16), ("stack", 0, 4),
           ("stack", 0, 1)]
IDS = [f"{k}{s or ''}{'_q%d' % w if w else ''}" for k, s, w in KERNELS]


def _clustered(n, seed, dtype):
    from cosmokdtree_amd.synthetic import clustered_periodic
    return clustered_periodic(n, 100.0, 0.5, seed=seed, dtype=dtype), np.full(3, 100.0)


CASES = {
    "clu3_f32_leaf40": lambda: (*_clustered(60000, 5, np.float32), 40),
    "clu3_f32_leaf16": lambda: (*_clustered(60000, 6, np.float32), 16),
    "u3_f64_leaf16": lambda: (np.random.default_rng(8).random((40000, 3)), None, 16),
    "p2_f32_leaf8": lambda: (((np.random.default_rng(9).random((30000, 2)) - 0.5) * 10).astype(np.float32),
                             np.full(2, 10.0), 8),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def kcase(request, ckd):
    pts, L, leaf = CASES[request.param]()
    print(f"\n[kernels {request.param}] n={len(pts)} leaf={leaf}", flush=True)
    tree = ckd.kdtree._build(pts, leaf, L)
    orc = OracleTree(pts, leaf, L)
    rng = np.random.default_rng(12)
    q = pts[rng.integers(0, len(pts), 6000)].astype(np.float64)
    q = q + (rng.random(q.shape) - 0.5) * (0.3 if L is not None else 0.003)
    if L is not None:
        q = np.clip(q, -L / 2, L / 2)
    q = q.astype(pts.dtype)
    yield dict(pts=pts, L=L, tree=tree, orc=orc, q=q)
    tree.knn_config("auto")
    tree.free()
    orc.close()


@pytest.mark.parametrize("kernel,stack,qpw", KERNELS, ids=IDS)
@pytest.mark.parametrize("k", [8, 32, 64])
def test_kernel_matches_oracle(kcase, kernel, stack, qpw, k):
    q = kcase["q"]
    d_ref, i_ref = kcase["orc"].knn(q, k, True)
    tree = kcase["tree"].knn_config(kernel, stack, qpw)
    st = torch.zeros(6, dtype=torch.int64, device="cuda")
    d, i = tree.knn(torch.from_numpy(q), k, sorted=True, stats=st)
    assert_knn_rows(kcase["pts"], q, k, d_ref, i_ref, d.cpu().numpy(), i.cpu().numpy(), L=kcase["L"])
    assert int(st[3]) == len(q)


@pytest.mark.parametrize("kernel,stack,qpw", KERNELS, ids=IDS)
def test_kernel_seeded(kcase, kernel, stack, qpw):
    """Seeds around the true k-th distance: about half are short and retried."""
    q = kcase["q"]
    k = 16
    tree = kcase["tree"].knn_config(kernel, stack, qpw)
    d0, i0 = tree.knn(torch.from_numpy(q), k, sorted=True)
    rng = np.random.default_rng(4)
    seed = (d0[:, k - 1].double().cpu().numpy() * rng.uniform(0.5, 1.5, len(q))).astype(np.float32)
    st = torch.zeros(6, dtype=torch.int64, device="cuda")
    d1, i1 = tree.knn(torch.from_numpy(q), k, sorted=True, seed_r=torch.from_numpy(seed), stats=st)
    assert torch.equal(d0, d1)
    assert_knn_rows(kcase["pts"], q, k, d0.cpu().numpy(), i0.cpu().numpy(), d1.cpu().numpy(), i1.cpu().numpy(),
                    L=kcase["L"])
    s = st.cpu().numpy()
    assert s[3] == len(q)
    assert 0.2 * len(q) < s[2] < 0.8 * len(q)  # retries happened, each counted once


def test_deep_tree_falls_back(ckd):
    """Exponentially spaced points make a tree deeper than 64 levels."""
    rng = np.random.default_rng(21)
    n = 400
    x = 1.25 ** -np.arange(n)
    pts = np.stack([x, rng.random(n) * 1e-60, rng.random(n) * 1e-60], axis=1)
    tree = ckd.kdtree._build(pts, 4, None)
    orc = OracleTree(pts, 4, None)
    try:
        assert tree.info().depth > 64
        q = np.concatenate([pts, pts * 1.0001], axis=0)
        q = np.concatenate([q] * 3, axis=0)  # several waves, input order
        d_ref, i_ref = orc.knn(q, 8, True)
        for kernel, stack, qpw in KERNELS:
            d, i = tree.knn_config(kernel, stack, qpw).knn(torch.from_numpy(q), 8, sorted=True)
            assert_knn_rows(pts, q, 8, d_ref, i_ref, d.cpu().numpy(), i.cpu().numpy())
    finally:
        tree.free()
        orc.close()


def test_knn_config_rejects(ckd):
    pts = np.random.default_rng(0).random((100, 3))
    tree = ckd.kdtree._build(pts, 16, None)
    try:
        with pytest.raises(ValueError):
            tree.knn_config("warp")
        with pytest.raises(Exception):
            tree.knn_config("stack", 48)
        for qpw in (3, 12, 128, -1):
            with pytest.raises(Exception):
                tree.knn_config("stack", 0, qpw)
    finally:
        tree.free()
