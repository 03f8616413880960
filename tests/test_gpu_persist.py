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


def test_load_recomputes_depth(ckd, tmp_path):
    """The traversals size their LDS stacks by the tree depth: a file whose
    header understates it loads with the depth its node links give, and
    answers as the built tree does; a node link out of range is refused."""
    import struct
    from cosmokdtree_amd.synthetic import clustered_periodic
    pts = clustered_periodic(40000, 100.0, 0.5, seed=19, dtype=np.float32)
    tree = ckd.kdtree._build(pts, 16, np.full(3, 100.0))
    path = tmp_path / "tree.ckd"
    tree.save(path)
    raw = bytearray(path.read_bytes())
    depth_at = 8 + 5 * 8 + 5 * 4  # PersistHeader: magic, 5 int64, then ndim..idx_bits, depth
    assert struct.unpack_from("<i", raw, depth_at)[0] == tree.info().depth
    struct.pack_into("<i", raw, depth_at, 1)
    bad = tmp_path / "shallow.ckd"
    bad.write_bytes(bytes(raw))
    back = ckd.kdtree.load_tree(bad)
    try:
        assert back.info().depth == tree.info().depth > 1
        q = torch.from_numpy(pts[::5]).cuda()
        d0, i0 = tree.knn(q, 32)
        d1, i1 = back.knn(q, 32)
        assert torch.equal(d0, d1) and torch.equal(i0, i1)
    finally:
        back.free()
    # first node's child pointed past the node array
    raw = bytearray(path.read_bytes())
    hdr = 8 + 5 * 8 + 6 * 4 + 8 * 8
    struct.pack_into("<i", raw, hdr + 8, 1 << 30)
    bad.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="child index"):
        ckd.kdtree.load_tree(bad)
    tree.free()
