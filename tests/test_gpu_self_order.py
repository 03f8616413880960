"""Queries that are the tree's own points (nq == n): the batch is taken in
the tree's row order, with each query's leaf read off its row, instead of
the locate pass and the counting sort.  Whether a batch is the tree's points
is decided on the device from every 16th row; a batch that is not (another
order, other points of the same count) is ordered from its coordinates, and
one that differs in a single row may go either way.  Every variant must give
the oracle's rows: the ordering only groups the walks and picks the rows
that prefill the lists."""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree
from tests.parity import assert_csr_equal, assert_knn_rows

pytestmark = pytest.mark.gpu


def _points(kind):
    rng = np.random.default_rng(5)
    if kind == "clustered_f32":
        from cosmokdtree_amd.synthetic import clustered_periodic
        return clustered_periodic(30000, 100.0, 0.5, seed=5, dtype=np.float32), np.full(3, 100.0)
    return rng.random((20000, 3)).astype(np.float64), None


def _variants(pts):
    rng = np.random.default_rng(9)
    one = pts.copy()
    one[len(pts) // 2] = pts[0]  # a duplicate of row 0 in place of one row
    other = pts[rng.integers(0, len(pts), len(pts))].astype(np.float64)
    other = (other + (rng.random(other.shape) - 0.5) * 0.01).astype(pts.dtype)
    return {"same": pts.copy(), "permuted": pts[rng.permutation(len(pts))], "one_row_changed": one,
            "other_points": other}


@pytest.fixture(scope="module", params=["clustered_f32", "uniform_f64"])
def built(request, ckd):
    pts, L = _points(request.param)
    print(f"\n[self-order {request.param}] n={len(pts)}", flush=True)
    tree = ckd.kdtree._build(pts, 16, L)
    orc = OracleTree(pts, 16, L)
    yield dict(pts=pts, L=L, tree=tree, orc=orc)
    tree.free()
    orc.close()


@pytest.mark.parametrize("variant", ["same", "permuted", "one_row_changed", "other_points"])
@pytest.mark.parametrize("k", [8, 32])
def test_knn_self_order(built, variant, k):
    q = _variants(built["pts"])[variant]
    d_ref, i_ref = built["orc"].knn(q, k, True)
    d_got, i_got = built["tree"].knn(torch.from_numpy(q), k, sorted=True)
    assert_knn_rows(built["pts"], q, k, d_ref, i_ref, d_got.cpu().numpy(), i_got.cpu().numpy(), L=built["L"])


def test_knn_self_order_same_tensor(built):
    """The very array the tree was built from, on the device, twice in a row
    (the second call reuses the scratch whose count the first one left)."""
    P = torch.from_numpy(built["pts"]).cuda()
    d_ref, i_ref = built["orc"].knn(built["pts"], 16, True)
    for _ in range(2):
        d_got, i_got = built["tree"].knn(P, 16, sorted=True)
        assert_knn_rows(built["pts"], built["pts"], 16, d_ref, i_ref, d_got.cpu().numpy(), i_got.cpu().numpy(),
                        L=built["L"])


@pytest.mark.parametrize("variant", ["same", "permuted"])
def test_ball_self_order(built, variant):
    q = _variants(built["pts"])[variant]
    r = 1.0 if built["L"] is not None else 0.02
    off_r, d_r, i_r = built["orc"].ball_csr(q, r, True)
    off, d, i = built["tree"].ball(torch.from_numpy(q), r, sorted=True)
    assert_csr_equal(off_r, d_r, i_r, off.cpu().numpy(), d.cpu().numpy(), i.cpu().numpy(), sorted_got=True)


def test_self_order_idx64(built, ckd):
    """The self-query ordering reads the stored original indices: with
    64-bit indices (LONGINT trees) the rows equal the 32-bit tree's and the
    oracle's."""
    pts, L = built["pts"], built["L"]
    t64 = ckd.kdtree._build(pts, 16, L, idx_bits=64)
    try:
        assert t64.info().idx_bits == 64
        P = torch.from_numpy(pts).cuda()
        d64, i64 = t64.knn(P, 32, sorted=True)
        d32, i32 = built["tree"].knn(P, 32, sorted=True)
        assert torch.equal(d32, d64) and torch.equal(i32, i64)
        d_ref, i_ref = built["orc"].knn(pts[::97], 32, True)
        assert_knn_rows(pts, pts[::97], 32, d_ref, i_ref, d64.cpu().numpy()[::97], i64.cpu().numpy()[::97], L=L)
    finally:
        t64.free()


@pytest.mark.parametrize("ndim,kernel", [(3, "lane"), (4, "auto"), (3, "auto")])
def test_self_order_points_on_split_planes(ckd, ndim, kernel):
    """Points on a 3^ndim lattice, dozens to hundreds at each site: the
    sliding midpoint peels one point of a site at a time into a leaf of its
    own, left of a split equal to its coordinate, while the descent by split
    values sends that point right.  The leaf holding a query's row is then
    not the leaf the per-query walk (which climbs from its leaf) must start
    from, so that walk takes the descent's leaves; the stack walk uses the
    row's leaf only to prefill.  Self-queries through the per-query kernel
    (forced, or the only one for ndim > 3) and through the default kernel
    must give the oracle's rows."""
    rng = np.random.default_rng(17)
    pts = rng.integers(0, 3, (5000, ndim)).astype(np.float64)
    tree = ckd.kdtree._build(pts, 16, None)
    orc = OracleTree(pts, 16, None)
    try:
        if kernel != "auto":
            tree.knn_config(kernel)
        d, i = tree.knn(torch.from_numpy(pts), 16, sorted=True)
        d_ref, i_ref = orc.knn(pts, 16, True)
        assert_knn_rows(pts, pts, 16, d_ref, i_ref, d.cpu().numpy(), i.cpu().numpy())
    finally:
        tree.free()
        orc.close()


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("kernel,qpw", [("auto", 0), ("stack", 0), ("stack", 4), ("wave", 0), ("lane", 0)])
def test_self_query_every_kernel(ckd, ndim, periodic, kernel, qpw):
    """Self-queries (the tree's own points, a fresh array) through every kNN
    kernel and periodicity, float32, with k inside one list and past it;
    coarse coordinates put many points on split planes."""
    rng = np.random.default_rng(100 + ndim)
    n = 6000
    L = np.full(ndim, 16.0) if periodic else None
    pts = (np.floor(rng.random((n, ndim)) * 64) / 4 - (8.0 if periodic else 0.0)).astype(np.float32)
    tree = ckd.kdtree._build(pts, 16, L)
    orc = OracleTree(pts, 16, L)
    try:
        tree.knn_config(kernel, 0, qpw)
        for k in (8, 40):
            d, i = tree.knn(torch.from_numpy(pts.copy()), k, sorted=True)
            d_ref, i_ref = orc.knn(pts, k, True)
            assert_knn_rows(pts, pts, k, d_ref, i_ref, d.cpu().numpy(), i.cpu().numpy(), L=L)
    finally:
        tree.free()
        orc.close()
