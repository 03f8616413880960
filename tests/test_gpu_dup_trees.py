"""GPU trees of inputs with many equal coordinates equal the committed
regression fixture node for node (tests/golden/make_gpu_dup_trees.py says
why these are pinned to the GPU build's own order), and their queries
match the oracle (kNN distances bit-exact with ties as sets, ball and box
sets), which holds whatever slide the build chose."""
import os

import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree
from tests.golden import make_gpu_dup_trees as gen
from tests.parity import assert_csr_equal, assert_knn_rows

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "gpu_dup_trees.npz")


@pytest.mark.parametrize("name", sorted(gen.cases()))
def test_dup_tree_matches_fixture(ckd, name):
    pts, leaf, L = gen.cases()[name]
    want = np.load(FIX)
    t = ckd.kdtree._build(pts, leaf, L)
    try:
        got = t.dump()
        for f in gen.FIELDS:
            w = want[f"{name}/{f}"]
            assert got[f].shape == w.shape, f
            assert np.array_equal(got[f].view(np.uint8), w.view(np.uint8)), f"{name}: {f} differs"
        orc = OracleTree(pts, leaf, L)
        try:
            q = pts[::7].astype(pts.dtype)
            d_ref, i_ref = orc.knn(q, 12, True)
            d, i = t.knn(torch.from_numpy(q), 12, sorted=True)
            assert_knn_rows(pts, q, 12, d_ref, i_ref, d.cpu().numpy(), i.cpu().numpy(), L=L)
            r = 0.6
            off_r, d_r, i_r = orc.ball_csr(q, r, True)
            off, bd, bi = t.ball(torch.from_numpy(q), r, sorted=True)
            assert_csr_equal(off_r, d_r, i_r, off.cpu().numpy(), bd.cpu().numpy(), bi.cpu().numpy(), sorted_got=True)
        finally:
            orc.close()
    finally:
        t.free()
