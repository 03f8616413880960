"""Boxes taken modulo a periodic tree's box (ckd_box_count_periodic /
ckd_box_fill_periodic), against the oracle's rule and brute force.

The reference's box search does not wrap (cosmokdtree.f90:1260); the
library's rule splits a box into the pieces of each axis modulo L, clipped
to [-L/2, L/2], and searches every sub-box with the reference's own rules
(inclusive point test, strict child visits, :1296-1325).  The GPU rows must
hold the oracle's points exactly (bit-exact index sets, row counts equal),
never a point twice; on float64 trees they also equal the brute-force
membership "some whole number of periods puts the coordinate in [lo, hi]".
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree
from tests.parity import assert_csr_equal
from tests.test_oracle_math import brute_wrap_box, straddling_boxes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("idx_bits", [32, 64])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_wrapped_boxes_match_oracle(ckd, dtype, idx_bits):
    from cosmokdtree_amd.synthetic import clustered_periodic
    L = 100.0
    pts = clustered_periodic(60000, L, 0.5, seed=31, dtype=dtype)
    pts[:300, 0] = -L / 2  # points on the faces
    pts[300:600, 1] = L / 2
    rng = np.random.default_rng(32)
    boxes = straddling_boxes(rng, 3000, L, dtype)
    print(f"\n[wrapped boxes {np.dtype(dtype).name} idx{idx_bits}] n={len(pts)} boxes={len(boxes)}", flush=True)
    tree = ckd.kdtree._build(pts, 16, np.full(3, L), idx_bits=idx_bits)
    orc = OracleTree(pts, 16, np.full(3, L))
    try:
        off_r, idx_r = orc.box_csr(boxes, wrap=True)
        off, idx = tree.box(torch.from_numpy(boxes), wrap=True)
        off, idx = off.cpu().numpy(), idx.cpu().numpy()
        assert_csr_equal(off_r, None, idx_r, off, None, idx)
        for b in range(len(boxes)):
            row = idx[off[b]:off[b + 1]]
            assert len(np.unique(row)) == len(row), f"box {b}: a point twice"
        if dtype == np.float64:
            for b in range(0, len(boxes), 7):
                assert np.array_equal(np.sort(idx[off[b]:off[b + 1]]), brute_wrap_box(pts, boxes[b], np.full(3, L)))
        straddle = np.mean(np.any(boxes[:, 0::2] < -L / 2, 1) | np.any(boxes[:, 1::2] > L / 2, 1))
        assert straddle > 0.25
        # without wrap: the reference's rule (no point outside [-L/2, L/2] is found)
        off_n, idx_n = tree.box(torch.from_numpy(boxes))
        off_nr, idx_nr = orc.box_csr(boxes)
        assert_csr_equal(off_nr, None, idx_nr, off_n.cpu().numpy(), None, idx_n.cpu().numpy())
    finally:
        tree.free()
        orc.close()


def test_wrapped_boxes_drop_in_and_errors(ckd):
    kd = ckd.kdtree
    rng = np.random.default_rng(5)
    pts = (rng.random((5000, 3)) - 0.5) * 10.0
    tree = kd.build_kdtree(pts, boxsize=[10.0] * 3)
    try:
        box = np.array([[4.0, 6.0, -1.0, 1.0, -6.0, -3.5]])
        off, idx = kd.box_search_batch(tree, box, wrap=True)
        assert np.array_equal(np.sort(idx), brute_wrap_box(pts, box[0], np.full(3, 10.0)))
        assert len(idx) > len(kd.box_search(tree, box[0]))  # the reference's rule misses the wrapped parts
    finally:
        kd.deallocate_kdtree(tree)
    open_tree = kd.build_kdtree(pts)
    try:
        with pytest.raises(ValueError, match="periodic"):
            kd.box_search_batch(open_tree, box, wrap=True)
    finally:
        kd.deallocate_kdtree(open_tree)
