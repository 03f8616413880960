"""The one-walk ball search (ckd_ball_search) against the oracle.

The walk stages its rows and moves them to their CSR places once the counts
are scanned; when the rows exceed the caller's capacity it writes only the
offsets and ckd_ball_fill completes the call with a second walk.  Covered:
float32/float64, periodic or not, sorted or not, radii that take whole
subtrees (range records), capacities below, at and above the total, a
capacity so small that the staging area overflows, query batches below and
above the leaf-ordering threshold, and a 4-D tree (no stack walk: counted
and filled in two walks inside the call).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree
from tests.parity import assert_csr_equal

pytestmark = pytest.mark.gpu


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _search(lib, tree, Q, r, cap, sorted_):
    nq = Q.shape[0]
    off = torch.empty(nq + 1, dtype=torch.int64, device="cuda")
    d = torch.full((max(cap, 1),), -7.0, dtype=tree.torch_dtype, device="cuda")
    i = torch.full((max(cap, 1),), -7, dtype=torch.int64, device="cuda")
    tot = C.c_int64(-1)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.ckd_ball_search(tree._ptr, _vp(Q), nq, float(r), int(sorted_), _vp(off), cap, _vp(d), _vp(i),
                             C.byref(tot), s)
    assert rc == 0, lib.ckd_last_error()
    tot = tot.value
    if tot > cap:  # rows not written: the second walk completes the call
        d = torch.empty(max(tot, 1), dtype=tree.torch_dtype, device="cuda")
        i = torch.empty(max(tot, 1), dtype=torch.int64, device="cuda")
        rc = lib.ckd_ball_fill(tree._ptr, _vp(Q), nq, float(r), int(sorted_), _vp(off), _vp(d), _vp(i), s)
        assert rc == 0, lib.ckd_last_error()
    torch.cuda.synchronize()
    return tot, off.cpu().numpy(), d[:tot].cpu().numpy(), i[:tot].cpu().numpy()


CASES = [
    # name, n, dtype, periodic, nq
    ("open_f32", 60000, np.float32, False, 20000),
    ("per_f32", 60000, np.float32, True, 20000),
    ("per_f64", 40000, np.float64, True, 9000),
    ("open_f64_small_batch", 40000, np.float64, False, 1000),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request, ckd):
    name, n, dt, periodic, nq = request.param
    rng = np.random.default_rng(5)
    from cosmokdtree_amd.synthetic import clustered_periodic
    pts = clustered_periodic(n, 20.0, 0.5, seed=9, dtype=dt)  # in [-10, 10): clustered
    L = np.full(3, 20.0) if periodic else None
    q = pts[rng.integers(0, n, nq)] + ((rng.random((nq, 3)) - 0.5) * 0.2).astype(dt)
    q = np.clip(q, -10, 10).astype(dt)
    print(f"\n[ball_search {name}] n={n} nq={nq} {np.dtype(dt).name} periodic={periodic}", flush=True)
    tree = ckd.kdtree._build(pts, 16, L)
    orc = OracleTree(pts, 16, L)
    from cosmokdtree_amd._lib import lib
    yield dict(pts=pts, q=q, tree=tree, orc=orc, lib=lib)
    tree.free()
    orc.close()


@pytest.mark.parametrize("r", [0.15, 0.6, 2.5])
@pytest.mark.parametrize("sorted_", [False, True])
def test_one_walk_matches_oracle(case, r, sorted_):
    q = case["q"]
    off_r, d_r, i_r = case["orc"].ball_csr(q, r, sorted_)
    Q = torch.from_numpy(q).cuda()
    total = int(off_r[-1])
    print(f"  r={r} total={total}", flush=True)
    for cap in (total + 1000, total):
        tot, off, d, i = _search(case["lib"], case["tree"], Q, r, cap, sorted_)
        assert tot == total
        assert_csr_equal(off_r, d_r, i_r, off, d, i, sorted_got=sorted_)


@pytest.mark.parametrize("shortfall", ["minus_one", "zero", "tiny"])
def test_capacity_short(case, shortfall):
    q = case["q"]
    r = 0.6
    off_r, d_r, i_r = case["orc"].ball_csr(q, r, True)
    total = int(off_r[-1])
    cap = {"minus_one": max(total - 1, 0), "zero": 0, "tiny": 1}[shortfall]
    tot, off, d, i = _search(case["lib"], case["tree"], torch.from_numpy(q).cuda(), r, cap, True)
    assert tot == total
    assert_csr_equal(off_r, d_r, i_r, off, d, i, sorted_got=True)


def test_device_api_hint_and_radius_changes(case):
    """tree.ball sizes its buffers from the previous call: a larger radius
    than before (rows beyond the guess) and a smaller one both stay exact."""
    q = case["q"]
    Q = torch.from_numpy(q).cuda()
    for r in (0.3, 1.5, 0.1, 1.5):
        off_r, d_r, i_r = case["orc"].ball_csr(q, r, False)
        off, d, i = case["tree"].ball(Q, r)
        assert_csr_equal(off_r, d_r, i_r, off.cpu().numpy(), d.cpu().numpy(), i.cpu().numpy())


def test_four_dimensions_two_walks(ckd):
    from cosmokdtree_amd._lib import lib
    rng = np.random.default_rng(3)
    pts = rng.random((8000, 4))
    q = rng.random((600, 4))
    tree = ckd.kdtree._build(pts, 16, None)
    orc = OracleTree(pts, 16, None)
    try:
        off_r, d_r, i_r = orc.ball_csr(q, 0.2, True)
        for cap in (int(off_r[-1]), 0):
            tot, off, d, i = _search(lib, tree, torch.from_numpy(q).cuda(), 0.2, cap, True)
            assert tot == int(off_r[-1])
            assert_csr_equal(off_r, d_r, i_r, off, d, i, sorted_got=True)
    finally:
        tree.free()
        orc.close()


def test_rejects_bad_arguments(ckd):
    from cosmokdtree_amd._lib import lib
    pts = np.random.default_rng(1).random((100, 3))
    tree = ckd.kdtree._build(pts, 16, None)
    try:
        Q = torch.from_numpy(pts[:10]).cuda()
        off = torch.empty(11, dtype=torch.int64, device="cuda")
        tot = C.c_int64(0)
        s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert lib.ckd_ball_search(tree._ptr, _vp(Q), 10, -1.0, 0, _vp(off), 0, None, None, C.byref(tot), s) != 0
        assert lib.ckd_ball_search(tree._ptr, _vp(Q), 10, 0.1, 0, _vp(off), 5, None, None, C.byref(tot), s) != 0
        assert lib.ckd_ball_search(tree._ptr, _vp(Q), 10, 0.1, 0, None, 0, None, None, C.byref(tot), s) != 0
    finally:
        tree.free()
