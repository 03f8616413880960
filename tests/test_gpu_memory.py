"""Device memory from the library's caching allocator: trees built one
after another (reusing freed blocks), on other streams, after the cache
has been released, and query workspaces that grow on a second stream all give
the results of the first tree, node for node and row for
row.  The oracle pins the first tree."""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree

pytestmark = pytest.mark.gpu


def _points():
    from cosmokdtree_amd.synthetic import clustered_periodic
    return clustered_periodic(40000, 100.0, 0.5, seed=31, dtype=np.float32), np.full(3, 100.0)


def _same_tree(a, b):
    da, db = a.dump(), b.dump()
    for key in da:
        assert np.array_equal(da[key], db[key]), key


def test_rebuilds_reuse_blocks(ckd):
    pts, L = _points()
    q = torch.from_numpy(pts[::5]).cuda()
    first = ckd.kdtree._build(pts, 16, L)
    orc = OracleTree(pts, 16, L)
    try:
        d0, i0 = first.knn(q, 16)
        d_ref, i_ref = orc.knn(q.cpu().numpy(), 16, True)
        np.testing.assert_array_equal(d0.cpu().numpy(), d_ref)
        np.testing.assert_array_equal(i0.cpu().numpy(), i_ref)
        for rep in range(3):  # freed blocks of the last build come back
            t = ckd.kdtree._build(pts, 16, L)
            try:
                _same_tree(first, t)
                d1, i1 = t.knn(q, 16)
                assert torch.equal(d0, d1) and torch.equal(i0, i1), rep
            finally:
                t.free()
        ckd.kdtree.release_memory()  # live trees keep their arrays
        d2, i2 = first.knn(q, 16)
        assert torch.equal(d0, d2) and torch.equal(i0, i2)
        t = ckd.kdtree._build(pts, 16, L)  # on an emptied cache
        try:
            _same_tree(first, t)
        finally:
            t.free()
    finally:
        first.free()
        orc.close()


def test_other_streams(ckd):
    pts, L = _points()
    q = torch.from_numpy(pts[::3]).cuda()
    ref = ckd.kdtree._build(pts, 16, L)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    try:
        d0, i0 = ref.knn(q[:500], 32)
        dl, il = ref.knn(q, 32)
        s1.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s1):
            t = ckd.kdtree._build(pts, 16, L)
            ds, is_ = t.knn(q[:500], 32)  # small workspaces on s1
        s1.synchronize()
        try:
            _same_tree(ref, t)
            assert torch.equal(ds, d0) and torch.equal(is_, i0)
            s2.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s2):
                dg, ig = t.knn(q, 32)  # the workspaces grow on s2
                off, bd, bi = t.ball(q[:3000], 0.5)
            s2.synchronize()
            assert torch.equal(dg, dl) and torch.equal(ig, il)
            o0, bd0, bi0 = ref.ball(q[:3000], 0.5)
            assert torch.equal(off, o0) and torch.equal(bi, bi0) and torch.equal(bd, bd0)
        finally:
            t.free()
    finally:
        ref.free()
