"""Pin the oracle by mathematics, independently of the reference.

* kNN equals brute force: the k smallest of all squared distances (computed
  in numpy in axis order, one rounding per operation), strict acceptance.
* ball equals brute force with the inclusive test d^2 <= r^2 (r cast to the
  tree precision first).
* box equals brute force with the inclusive per-axis test, except where the
  reference's strict child-visit rule drops a point lying exactly on a split
  that is also the box edge (cosmokdtree.f90:1296-1302); random boxes never
  hit that case.
* The tree obeys the sliding-midpoint invariants (cosmokdtree.f90:347-422).
* A mistake planted in the oracle (a swapped axis, a wrong sign) must make
  these checks fail, which shows the checks can see such mistakes.
"""
import numpy as np
import pytest

from oracle.oracle import OracleTree
from tests.parity import assert_knn_rows, ref_dist2


def brute_knn(pts, q, k, L=None):
    nq = len(q)
    d = np.empty((nq, k), pts.dtype)
    i = np.empty((nq, k), np.int64)
    for r in range(nq):
        d2 = ref_dist2(pts, q[r], L)
        o = np.lexsort((np.arange(len(pts)), d2))[:k]
        dd = np.sqrt(d2[o])
        ii = o.astype(np.int64)
        if len(o) < k:
            pad = k - len(o)
            dd = np.concatenate([dd, np.full(pad, np.sqrt(pts.dtype.type(np.finfo(np.float32).max)))])
            ii = np.concatenate([ii, np.full(pad, -2)])
        d[r], i[r] = dd, ii
    return d, i


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("periodic", [False, True])
def test_knn_matches_brute_force(dtype, periodic):
    rng = np.random.default_rng(3)
    if periodic:
        L = np.full(3, 10.0)
        pts = ((rng.random((3000, 3)) - 0.5) * 10).astype(dtype)
        q = ((rng.random((200, 3)) - 0.5) * 10).astype(dtype)
    else:
        L = None
        pts = rng.random((3000, 3)).astype(dtype)
        q = rng.random((200, 3)).astype(dtype)
    t = OracleTree(pts, 8, L)
    for k in (1, 7, 40):
        d, i = t.knn(q, k, True)
        bd, bi = brute_knn(pts, q, k, L)
        assert_knn_rows(pts, q, k, bd, bi, d, i, L=L, sorted_got=False)
        assert np.all(np.diff(d, axis=1) >= 0)


def test_knn_more_than_n():
    rng = np.random.default_rng(4)
    pts = rng.random((7, 3))
    q = rng.random((3, 3))
    t = OracleTree(pts, 2)
    d, i = t.knn(q, 10, True)
    bd, bi = brute_knn(pts, q, 10)
    assert np.array_equal(d, bd)
    assert np.all(i[:, 7:] == -2)
    assert np.array_equal(np.sort(i[:, :7], axis=1), np.tile(np.arange(7), (3, 1)))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("periodic", [False, True])
def test_ball_matches_brute_force(dtype, periodic):
    rng = np.random.default_rng(5)
    L = np.full(3, 10.0) if periodic else None
    scale = 10.0 if periodic else 1.0
    pts = ((rng.random((4000, 3)) - (0.5 if periodic else 0)) * scale).astype(dtype)
    q = ((rng.random((150, 3)) - (0.5 if periodic else 0)) * scale).astype(dtype)
    r = 0.9 if periodic else 0.09
    t = OracleTree(pts, 16, L)
    off, d, i = t.ball_csr(q, r, True)
    r2 = dtype(r) * dtype(r)
    for j in range(len(q)):
        d2 = ref_dist2(pts, q[j], L)
        want = np.nonzero(d2 <= r2)[0]
        got = i[off[j]:off[j + 1]]
        assert np.array_equal(np.sort(got), want)
        assert np.array_equal(np.sort(d[off[j]:off[j + 1]]), np.sort(np.sqrt(d2[want])))


def test_box_matches_brute_force():
    rng = np.random.default_rng(6)
    pts = rng.random((5000, 3))
    t = OracleTree(pts, 16)
    c = rng.random((100, 3))
    h = rng.random((100, 3)) * 0.1
    boxes = np.empty((100, 6))
    boxes[:, 0::2] = c - h
    boxes[:, 1::2] = c + h
    off, i = t.box_csr(boxes)
    for j in range(100):
        inside = np.all((pts >= boxes[j, 0::2]) & (pts <= boxes[j, 1::2]), axis=1)
        assert np.array_equal(np.sort(i[off[j]:off[j + 1]]), np.nonzero(inside)[0])


def check_invariants(pts, dump, leaf, L=None):
    """Sliding-midpoint invariants of a preorder dump."""
    n, D = pts.shape
    lpts = dump["lpts"]
    lidx = dump["lidx"]
    assert np.array_equal(np.sort(lidx), np.arange(n))  # every point in exactly one leaf
    assert np.array_equal(lpts, pts[lidx])
    if L is None:
        assert np.array_equal(dump["minb"][0], pts.min(axis=0))
        assert np.array_equal(dump["maxb"][0], pts.max(axis=0))
    else:
        assert np.all(dump["minb"][0] == -np.asarray(L, pts.dtype) / 2)
    for v in range(len(dump["axis"])):
        s, c = dump["start"][v], dump["count"][v]
        P = lpts[s:s + c]
        assert np.all(P >= dump["minb"][v]) and np.all(P <= dump["maxb"][v])  # points lie in their cell
        if dump["is_leaf"][v]:
            assert 1 <= c <= leaf
            continue
        assert c > leaf
        ax, sp = dump["axis"][v], dump["split"][v]
        side = dump["maxb"][v] - dump["minb"][v]
        assert side[ax] == side.max()  # longest cell side
        l, r = dump["left"][v], dump["right"][v]
        Pl = lpts[dump["start"][l]:dump["start"][l] + dump["count"][l]]
        Pr = lpts[dump["start"][r]:dump["start"][r] + dump["count"][r]]
        assert dump["count"][l] + dump["count"][r] == c
        mid = pts.dtype.type(0.5) * (dump["minb"][v][ax] + dump["maxb"][v][ax])
        mn, mx = P[:, ax].min(), P[:, ax].max()
        if mn >= mid:  # slide to the minimum: one point on the left
            assert sp == mn and dump["count"][l] == 1 and Pl[0, ax] == mn
        elif mx < mid:  # slide to the maximum: one point on the right
            assert sp == mx and dump["count"][r] == 1 and Pr[0, ax] == mx
        else:
            assert sp == mid and np.all(Pl[:, ax] < sp) and np.all(Pr[:, ax] >= sp)
        assert dump["maxb"][l][ax] == sp and dump["minb"][r][ax] == sp


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("leaf", [1, 16])
def test_tree_invariants(periodic, leaf):
    rng = np.random.default_rng(8)
    if periodic:
        L = np.full(3, 8.0)
        pts = np.concatenate([(rng.random((1500, 3)) - 0.5) * 8, 0.05 * rng.standard_normal((500, 3))])
    else:
        L = None
        pts = np.concatenate([rng.random((1500, 3)), 0.3 + 0.01 * rng.standard_normal((500, 3))])
    t = OracleTree(pts, leaf, L)
    check_invariants(pts, t.dump(), leaf, L)


def test_planted_mistakes_are_caught():
    """Plant a wrong sign and a transposed operand in the distance and check
    that the brute-force comparison rejects the planted oracle."""
    rng = np.random.default_rng(9)
    pts = rng.random((500, 3))
    q = rng.random((20, 3))
    t = OracleTree(pts, 8)
    d, i = t.knn(q, 5, True)
    bd, bi = brute_knn(pts, q, 5)

    def planted_dist2(points, qq, L=None):  # axis 1 and 2 swapped in the query
        return ref_dist2(points, qq[[0, 2, 1]], L)

    wd, wi = [], []
    for r in range(len(q)):
        d2 = planted_dist2(pts, q[r])
        o = np.argsort(d2, kind="stable")[:5]
        wd.append(np.sqrt(d2[o]))
        wi.append(o)
    with pytest.raises(AssertionError):
        assert_knn_rows(pts, q, 5, np.array(wd), np.array(wi), d, i, sorted_got=False)
    # a dropped term: squared distance without the last axis
    two = [np.sqrt(ref_dist2(pts[:, :2], q[r, :2])) for r in range(len(q))]
    wd = np.array([np.sort(x)[:5] for x in two])
    wi = np.array([np.argsort(x, kind="stable")[:5] for x in two])
    with pytest.raises(AssertionError):
        assert_knn_rows(pts, q, 5, wd, wi, d, i, sorted_got=False)
    assert_knn_rows(pts, q, 5, bd, bi, d, i, sorted_got=False)
