"""Queries that overlap on one tree: the tree is read-only and every call
takes its own scratch, as the reference's tree serves concurrent OpenMP
queries (src/cosmokdtree2py.f90:132-141).  Two streams, and two host
threads, query one tree at once; every result must equal the same query
run alone."""
import threading

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _tree_and_queries(ckd):
    from cosmokdtree_amd.synthetic import clustered_periodic
    pts = clustered_periodic(200000, 100.0, 0.5, seed=41, dtype=np.float32)
    tree = ckd.kdtree._build(pts, 16, np.full(3, 100.0))
    P = torch.from_numpy(pts).cuda()
    # two batches large enough to be leaf-ordered (scratch per call), of different sizes
    return tree, P[:60000].contiguous(), P[70000:150000].contiguous()


def test_two_streams_one_tree(ckd):
    tree, Q1, Q2 = _tree_and_queries(ckd)
    try:
        want1 = tree.knn(Q1, 32)
        want2 = tree.knn(Q2, 16)
        wball = tree.ball(Q2[:20000], 0.5)
        torch.cuda.synchronize()
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        s1.wait_stream(torch.cuda.current_stream())
        s2.wait_stream(torch.cuda.current_stream())
        got = []
        for rep in range(3):
            with torch.cuda.stream(s1):
                a = tree.knn(Q1, 32)
            with torch.cuda.stream(s2):
                b = tree.knn(Q2, 16)
            with torch.cuda.stream(s1):
                c = tree.ball(Q2[:20000], 0.5)
            got.append((a, b, c))
        torch.cuda.synchronize()
        for a, b, c in got:
            assert torch.equal(a[0], want1[0]) and torch.equal(a[1], want1[1])
            assert torch.equal(b[0], want2[0]) and torch.equal(b[1], want2[1])
            assert torch.equal(c[0], wball[0]) and torch.equal(torch.sort(c[2])[0], torch.sort(wball[2])[0])
    finally:
        tree.free()


def test_two_host_threads_one_tree(ckd):
    tree, Q1, Q2 = _tree_and_queries(ckd)
    try:
        want = {32: tree.knn(Q1, 32), 16: tree.knn(Q2, 16)}
        torch.cuda.synchronize()
        out, errs = {}, []

        def run(k, Q):
            try:
                s = torch.cuda.Stream()
                with torch.cuda.stream(s):
                    res = [tree.knn(Q, k) for _ in range(4)]
                s.synchronize()
                out[k] = res
            except Exception as e:  # reported below
                errs.append(e)

        ths = [threading.Thread(target=run, args=(32, Q1)), threading.Thread(target=run, args=(16, Q2))]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        assert not errs, errs
        for k, res in out.items():
            for d, i in res:
                assert torch.equal(d, want[k][0]) and torch.equal(i, want[k][1])
    finally:
        tree.free()
