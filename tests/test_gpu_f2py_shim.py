"""The f2py-level interface (cosmokdtree_amd.pycosmokdtree) against the
drop-in API and the oracle.

The calls are made the way the reference's Python wrapper makes them
(src/cosmokdtree/kdtree.py:16-87): Fortran-ordered float64 points, the
(tree, target, ndim, ntar, k, sorted) kNN call and its 1-based indices,
the two-call ball and box protocols with zero-padded ball rows.
"""
import numpy as np
import pytest

from oracle.oracle import OracleTree
from tests.parity import assert_csr_equal, assert_knn_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shim(ckd):
    from cosmokdtree_amd import pycosmokdtree
    return pycosmokdtree.pycosmokdtree


@pytest.mark.parametrize("periodic", [False, True])
def test_shim_matches_oracle(shim, periodic):
    rng = np.random.default_rng(3)
    n, d = 20000, 3
    if periodic:
        L = np.full(d, 10.0)
        pts = (rng.random((n, d)) - 0.5) * 10.0
        boxsize = L
    else:
        L = None
        pts = rng.random((n, d))
        boxsize = -1.0 * np.ones(d)
    tree = shim.pybuild_kdtree(np.asfortranarray(pts), 16, np.asarray(boxsize, dtype=np.float64))
    assert isinstance(tree, int)
    orc = OracleTree(pts, 16, L)
    try:
        q = pts[rng.integers(0, n, 500)] + (rng.random((500, d)) - 0.5) * 0.01
        if periodic:
            q = np.clip(q, -5.0, 5.0)
        # kNN: 1-based indices, -1 for an empty slot
        for k in (8, 32):
            dist, idx = shim.pyknn_search(tree, q, d, len(q), k, True)
            assert dist.flags.f_contiguous and idx.dtype == np.int64 and idx.min() >= 1
            d_ref, i_ref = orc.knn(q, k, True)
            assert_knn_rows(pts, q, k, d_ref, i_ref, dist, idx - 1, L=L)
        # ball: counts + opaque handles, then zero-padded rows
        r = 0.3 if periodic else 0.03
        nball, nballmax, query = shim.pyball_search_call_1(tree, q, d, len(q), r, False)
        assert nball.dtype == np.int32 and len(query) == len(q)
        bdist, bidx = shim.pyball_search_call_2(d, len(q), query, nballmax, nball)
        assert bdist.shape == (len(q), nballmax) and bidx.shape == (len(q), nballmax)
        off = np.concatenate([[0], np.cumsum(nball)]).astype(np.int64)
        got_d = np.concatenate([bdist[j, :nball[j]] for j in range(len(q))])
        got_i = np.concatenate([bidx[j, :nball[j]] for j in range(len(q))]) - 1
        o_ref, d_ref, i_ref = orc.ball_csr(q, r, False)
        assert_csr_equal(o_ref, d_ref, i_ref, off, got_d, got_i)
        assert np.all(bidx[np.arange(nballmax)[None, :] >= nball[:, None]] == 0)  # zero padding
        # box: count + handle, then 1-based indices
        for j in range(5):
            c = q[j]
            box = np.ravel(np.stack([c - 0.05, c + 0.05], axis=1))  # (xmin, xmax, ymin, ...)
            nbox, handle = shim.pybox_search_call_1(tree, box)
            bi = shim.pybox_search_call_2(handle, nbox)
            ob, ib = orc.box_csr(box.reshape(1, -1))
            assert set((bi - 1).tolist()) == set(ib.tolist()) and nbox == len(ib)
    finally:
        shim.pydeallocate_tree(tree)
        orc.close()
