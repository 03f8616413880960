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


# ---------------------------------------------------- periodic (wrapped) boxes

def brute_wrap_box(pts, box, L):
    """Indices of the points inside a box taken modulo the periodic box:
    on every axis some whole number of periods j puts the coordinate in
    [lo, hi] (inclusive), in float64."""
    p = pts.astype(np.float64)
    keep = np.ones(len(p), dtype=bool)
    for d in range(p.shape[1]):
        lo, hi, Ld = float(box[2 * d]), float(box[2 * d + 1]), float(L[d])
        if hi - lo >= Ld:
            continue
        js = np.arange(np.floor((lo - 0.5 * Ld) / Ld) - 1, np.ceil((hi + 0.5 * Ld) / Ld) + 2)
        x = p[:, d][:, None] + js[None, :] * Ld
        keep &= np.any((x >= lo) & (x <= hi), axis=1)
    return np.nonzero(keep)[0]


def near_wrapped_edge(pts, box, L, tol):
    """Points within tol of a box face, modulo L (where float32 rounding of
    the shifted faces may decide differently from exact arithmetic)."""
    p = pts.astype(np.float64)
    near = np.zeros(len(p), dtype=bool)
    for d in range(p.shape[1]):
        for e in (float(box[2 * d]), float(box[2 * d + 1])):
            r = np.mod(p[:, d] - e + 0.5 * L[d], L[d]) - 0.5 * L[d]
            near |= np.abs(r) < tol
    return near


def straddling_boxes(rng, nb, L, dtype, D=3):
    """Boxes of side log-uniform in [0.5, 5] (times L/500) centred anywhere in
    the box, many crossing a face; a few wider than L on an axis, a few far
    outside [-L/2, L/2] and a few with lo > hi."""
    c = (rng.random((nb, D)) - 0.5) * L
    c[: nb // 3] = np.sign(c[: nb // 3]) * (0.5 * L - rng.random((nb // 3, D)) * 0.01 * L)  # near the faces
    side = np.exp(rng.uniform(np.log(0.5), np.log(5.0), (nb, D))) * L / 500.0 * 40
    b = np.empty((nb, 2 * D))
    b[:, 0::2] = c - side / 2
    b[:, 1::2] = c + side / 2
    b[-4, 0:2] = [-0.7 * L, 0.8 * L]  # wider than L on axis 0
    b[-3, :] += 2 * L                  # two periods away
    b[-2, :] -= 3 * L
    b[-1, 2:4] = [1.0, -1.0]           # empty on axis 1
    return b.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_wrapped_box_matches_brute_force(dtype):
    rng = np.random.default_rng(4242)
    L = np.full(3, 100.0)
    pts = ((rng.random((20000, 3)) - 0.5) * 100.0).astype(dtype)
    pts[:200, 0] = -50.0  # points on the faces
    pts[200:400, 2] = 50.0
    t = OracleTree(pts, 16, L)
    boxes = straddling_boxes(rng, 300, 100.0, dtype)
    off, idx = t.box_csr(boxes, wrap=True)
    crossing = 0
    for b in range(len(boxes)):
        got = idx[off[b]:off[b + 1]]
        assert len(np.unique(got)) == len(got), f"box {b}: a point twice"
        want = brute_wrap_box(pts, boxes[b], L)
        if dtype == np.float64:
            assert np.array_equal(np.sort(got), want), f"box {b}"
        else:
            diff = np.setxor1d(got, want)
            assert np.all(near_wrapped_edge(pts[diff], boxes[b], L, 1e-4)), f"box {b}: {diff[:5]}"
        crossing += bool(np.any(boxes[b, 0::2] < -50) or np.any(boxes[b, 1::2] > 50))
    assert crossing > 50
    # a box inside [-L/2, L/2] gives what the reference's box search gives
    inner = np.array([[-10.0, 5.0, -3.0, 20.0, 0.0, 40.0]], dtype=dtype)
    assert np.array_equal(t.box_csr(inner, wrap=True)[1], t.box_csr(inner)[1])


def test_wrapped_box_planted_mistakes_are_caught():
    """A rule that forgets the wrap on one axis, or shifts the wrong way,
    disagrees with the brute-force membership."""
    rng = np.random.default_rng(99)
    L = np.full(3, 10.0)
    pts = (rng.random((5000, 3)) - 0.5) * 10.0
    box = np.array([4.0, 6.0, -1.0, 1.0, -6.0, -3.5])  # crosses +L/2 on x and -L/2 on z
    want = brute_wrap_box(pts, box, L)
    # no wrap on z: clip only
    clipped = box.copy()
    clipped[4] = -5.0
    t = OracleTree(pts, 16, L)
    off, got = t.box_csr(np.array([[4.0, 5.0, -1.0, 1.0, -5.0, -3.5]]))
    assert not np.array_equal(np.sort(got), want)
    # shifted the wrong way on x: [4 + 10, 6 + 10] misses everything
    assert len(brute_wrap_box(pts, box, L)) > len(t.box_csr(np.array([[-5.0, -4.0, -1.0, 1.0, -5.0, -3.5]]))[1])
    assert np.array_equal(np.sort(t.box_csr(box[None], wrap=True)[1]), want)
