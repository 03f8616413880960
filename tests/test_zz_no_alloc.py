"""ckd_knn_noalloc (scratch in a static device array, nothing allocated)
against ckd_knn, through the C ABI: two calls in a row, the second with a
smaller batch, plain and seeded, and a batch beyond the static scratch
refused before any work.  Calls share the GPU's array, so they run one after
another here."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _noalloc(lib, tree, Q, k, sorted=True, seed=None, stats=None):
    nq = Q.shape[0]
    d = torch.full((nq, k), -1.0, dtype=Q.dtype, device=Q.device)
    i = torch.full((nq, k), -7, dtype=torch.int64, device=Q.device)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.ckd_knn_noalloc(tree._ptr, _vp(Q), nq, k, int(sorted), _vp(d), _vp(i), _vp(seed), _vp(stats), s)
    return rc, d, i


def test_noalloc_matches_knn(ckd):
    from cosmokdtree_amd._lib import lib
    from cosmokdtree_amd.synthetic import clustered_periodic
    pts = clustered_periodic(150000, 100.0, 0.5, seed=51, dtype=np.float32)
    tree = ckd.kdtree._build(pts, 16, np.full(3, 100.0))
    P = torch.from_numpy(pts).cuda()
    try:
        limit = 64 << 20
        for nq, k in ((60000, 32), (9000, 16), (3000, 8)):  # ordered, ordered and smaller, unordered
            Q = P[:nq].contiguous() + 0.001
            need = lib.ckd_knn_noalloc_scratch_bytes(tree._ptr, nq, 0)
            assert 0 < need <= limit if nq >= 4096 else need == 0
            print(f"\n[noalloc] nq={nq} k={k} scratch {need} B", flush=True)
            rc, d, i = _noalloc(lib, tree, Q, k)
            assert rc == 0, lib.ckd_last_error()
            wd, wi = tree.knn(Q, k)
            torch.cuda.synchronize()
            assert torch.equal(d, wd) and torch.equal(i, wi)
        # seeded, with seeds too short for some queries (the re-search lists in the static scratch)
        for nq in (40000, 12000):
            Q = P[-nq:].contiguous()
            seed = torch.full((nq,), 0.05, dtype=torch.float32, device="cuda")
            st = torch.zeros(6, dtype=torch.int64, device="cuda")
            rc, d, i = _noalloc(lib, tree, Q, 32, seed=seed, stats=st)
            assert rc == 0, lib.ckd_last_error()
            wd, wi = tree.knn(Q, 32)
            torch.cuda.synchronize()
            assert torch.equal(d, wd) and torch.equal(i, wi)
            assert int(st[2]) > 0  # some seeds were short
        # beyond the static scratch: refused, outputs untouched
        nq = limit // 12 + 4096
        Q = P[torch.arange(nq, device="cuda") % P.shape[0]].contiguous()
        assert lib.ckd_knn_noalloc_scratch_bytes(tree._ptr, nq, 0) > limit
        rc, d, i = _noalloc(lib, tree, Q, 1)
        torch.cuda.synchronize()
        assert rc == -5 and b"static scratch" in lib.ckd_last_error()
        assert bool(torch.all(i == -7))
        rc, d, i = _noalloc(lib, tree, Q[:100], 65)
        assert rc == -5
    finally:
        tree.free()
