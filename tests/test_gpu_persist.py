"""A saved and reloaded tree (or forest) answers every query exactly as the
one that was built."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_tree_roundtrip(ckd, tmp_path):
    from cosmokdtree_amd.synthetic import clustered_periodic
    pts = clustered_periodic(50000, 100.0, 0.5, seed=9, dtype=np.float32)
    tree = ckd.kdtree._build(pts, 16, np.full(3, 100.0))
    path = tmp_path / "tree.ckd"
    tree.save(path)
    back = ckd.kdtree.load_tree(path)
    try:
        a, b = tree.info(), back.info()
        for f in ("n", "ndim", "prec", "leafsize", "periodic", "idx_bits", "depth", "nnodes", "nleaves", "nsets"):
            assert getattr(a, f) == getattr(b, f), f
        da, db = tree.dump(), back.dump()
        for key in da:
            assert np.array_equal(da[key], db[key]), key
        q = torch.from_numpy(pts[::7]).cuda()
        for k in (8, 32):
            d0, i0 = tree.knn(q, k)
            d1, i1 = back.knn(q, k)
            assert torch.equal(d0, d1) and torch.equal(i0, i1)
        o0, bd0, bi0 = tree.ball(q[:2000], 1.0)
        o1, bd1, bi1 = back.ball(q[:2000], 1.0)
        assert torch.equal(o0, o1) and torch.equal(bi0, bi1) and torch.equal(bd0, bd1)
        boxes = torch.cat([q[:500] - 0.8, q[:500] + 0.8], 1)[:, [0, 3, 1, 4, 2, 5]].contiguous()
        ob0, ib0 = tree.box(boxes)
        ob1, ib1 = back.box(boxes)
        assert torch.equal(ob0, ob1) and torch.equal(ib0, ib1)
    finally:
        tree.free()
        back.free()


def test_forest_roundtrip(ckd, tmp_path):
    from cosmokdtree_amd import batch
    from cosmokdtree_amd.synthetic import halo_set
    sets = [halo_set(n, 10.0, seed=40 + i) for i, n in enumerate([30, 500, 4000, 12000])]
    forest = batch.build_forest(sets, leaf=16, boxsize=[10.0] * 3)
    path = tmp_path / "forest.ckd"
    forest.save(path)
    back = ckd.kdtree.load_tree(path)
    try:
        assert back.info().nsets == 4
        nq = [max(1, len(s) // 10) for s in sets]
        q = torch.from_numpy(np.concatenate([s[:m] for s, m in zip(sets, nq)])).cuda()
        ks = np.array([8, 16, 32, 64], np.int32)
        r = np.array([0.1, 0.2, 0.3, 0.4])
        out0 = batch.BatchPlan(forest, nq, ks, r).run(q)
        out1 = batch.BatchPlan(back, nq, ks, r).run(q)
        for key in ("knn_dist", "knn_idx", "ball_off"):
            assert torch.equal(out0[key], out1[key]), key
    finally:
        forest.free()
        back.free()
