"""Trees that store 64-bit original indices (the reference's LONGINT=1,
src/Makefile:6-31; ckd_build with idx_bits = 64) give the same results as
32-bit ones for every query kind, and the oracle's.  The index width only
changes how leaf rows map back to original indices (load_idx)."""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree
from tests.parity import assert_knn_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("periodic", [True, False])
def test_idx64_matches_idx32(ckd, periodic):
    from cosmokdtree_amd.synthetic import clustered_periodic
    rng = np.random.default_rng(41)
    if periodic:
        pts = clustered_periodic(80000, 100.0, 0.5, seed=41, dtype=np.float32)
        L = np.full(3, 100.0)
    else:
        pts = rng.random((80000, 3)).astype(np.float32)
        L = None
    t32 = ckd.kdtree._build(pts, 16, L)
    t64 = ckd.kdtree._build(pts, 16, L, idx_bits=64)
    orc = OracleTree(pts, 16, L)
    try:
        assert t32.info().idx_bits == 32 and t64.info().idx_bits == 64
        q = torch.from_numpy(pts[rng.integers(0, len(pts), 8000)])  # dense enough for leaf order + prefill
        for k in (8, 32, 64):
            d32, i32 = t32.knn(q, k)
            d64, i64 = t64.knn(q, k)
            assert torch.equal(d32, d64) and torch.equal(i32, i64), k
        d_ref, i_ref = orc.knn(q.numpy(), 32, True)
        d64, i64 = t64.knn(q, 32)
        assert_knn_rows(pts, q.numpy(), 32, d_ref, i_ref, d64.cpu().numpy(), i64.cpu().numpy(), L=L)
        r = 0.5 if periodic else 0.02
        o32, bd32, bi32 = t32.ball(q[:2000], r, sorted=True)
        o64, bd64, bi64 = t64.ball(q[:2000], r, sorted=True)
        assert torch.equal(o32, o64) and torch.equal(bd32, bd64) and torch.equal(bi32, bi64)
        c = q[:50].double().numpy()
        boxes = np.concatenate([c - 1.0, c + 1.0], axis=1)[:, [0, 3, 1, 4, 2, 5]] if periodic else \
            np.concatenate([c - 0.02, c + 0.02], axis=1)[:, [0, 3, 1, 4, 2, 5]]
        ob32, bx32 = t32.box(boxes)
        ob64, bx64 = t64.box(boxes)
        assert torch.equal(ob32, ob64)
        for j in range(len(boxes)):
            a, b = int(ob32[j]), int(ob32[j + 1])
            assert sorted(bx32[a:b].tolist()) == sorted(bx64[a:b].tolist())
        if periodic:
            h32, r32 = t32.sph_density(q, 32)
            h64, r64 = t64.sph_density(q, 32)
            assert torch.equal(h32, h64) and torch.equal(r32, r64)
    finally:
        t32.free()
        t64.free()
        orc.close()
