"""Parity checks between the GPU library and the oracle.

The bar (north star): indices and result sets bit-exact, distances
bit-identical in the tree precision, points tied at the k-th distance
compared as sets.  All distance recomputation here follows distance() of
the reference (cosmokdtree.f90:1441-1465) in numpy, whose float32/float64
elementwise operations round once per operation.
"""
from __future__ import annotations

import numpy as np


def ref_dist2(points, q, L=None):
    """Squared distances of all points to q, in the points' precision,
    summed in axis order from 0 (cosmokdtree.f90:1449-1462)."""
    dx = points - q.astype(points.dtype)
    if L is not None:
        a = np.abs(dx)
        b = np.asarray(L, dtype=points.dtype) - a
        dx = np.where(b < a, b, a)
    s = np.zeros(points.shape[0], dtype=points.dtype)
    for i in range(points.shape[1]):
        s = s + dx[:, i] * dx[:, i]
    return s


def assert_knn_rows(points, queries, k, d_ref, i_ref, d_got, i_got, L=None, sorted_got=True, rows=None):
    """Compare kNN results row by row.

    Distances must agree bit for bit as sorted multisets; indices must agree
    as sets below the k-th distance; entries at the k-th distance must be
    points that really are at that distance (the choice among tied points
    depends on visit order, cosmokdtree.f90:658).
    """
    d_ref = np.asarray(d_ref)
    d_got = np.asarray(d_got)
    i_ref = np.asarray(i_ref)
    i_got = np.asarray(i_got)
    assert d_got.shape == d_ref.shape and i_got.shape == i_ref.shape
    assert d_got.dtype == d_ref.dtype, (d_got.dtype, d_ref.dtype)
    nrows = d_ref.shape[0]
    rows = range(nrows) if rows is None else rows
    for r in rows:
        dr, ir, dg, ig = d_ref[r], i_ref[r], d_got[r], i_got[r]
        if sorted_got:
            assert np.all(dg[1:] >= dg[:-1]), f"row {r}: not sorted"
            tie = (dg[1:] == dg[:-1]) & (ig[1:] >= 0) & (ig[:-1] >= 0)
            assert np.all(ig[1:][tie] > ig[:-1][tie]), f"row {r}: equal distances not in index order"
        sr = np.sort(dr)
        sg = np.sort(dg)
        if not np.array_equal(sr.view(np.uint8), sg.view(np.uint8)):
            bad = np.nonzero(sr != sg)[0]
            raise AssertionError(f"row {r}: distances differ at {bad[:5]}: ref {sr[bad[:5]]} got {sg[bad[:5]]}")
        dmax = sr[-1]
        below_r = set(ir[dr < dmax].tolist())
        below_g = set(ig[dg < dmax].tolist())
        assert below_r == below_g, f"row {r}: index sets differ: {sorted(below_r ^ below_g)[:10]}"
        tie_g = ig[dg == dmax]
        tie_r = ir[dr == dmax]
        if np.array_equal(np.sort(tie_g), np.sort(tie_r)):
            continue
        # tied at the k-th distance: every chosen point must be at that distance
        valid = tie_g[tie_g >= 0]
        assert len(set(valid.tolist())) == len(valid), f"row {r}: duplicate indices"
        d2 = ref_dist2(points[valid], queries[r], L)
        d = np.sqrt(d2)
        assert np.array_equal(d, np.full_like(d, dmax)), f"row {r}: tied entries not at the k-th distance"


def assert_csr_equal(off_ref, d_ref, i_ref, off_got, d_got, i_got, sorted_got=False):
    """Ball/box results: per row the same (index, distance) pairs."""
    off_ref = np.asarray(off_ref)
    off_got = np.asarray(off_got)
    assert np.array_equal(off_ref, off_got), "row counts differ: " + str(
        np.nonzero(np.diff(off_ref) != np.diff(off_got))[0][:10])
    for r in range(len(off_ref) - 1):
        a, b = off_ref[r], off_ref[r + 1]
        ir, ig = np.asarray(i_ref[a:b]), np.asarray(i_got[a:b])
        oref = np.argsort(ir, kind="stable")
        ogot = np.argsort(ig, kind="stable")
        assert np.array_equal(ir[oref], ig[ogot]), f"row {r}: index sets differ"
        if d_ref is not None:
            dr, dg = np.asarray(d_ref[a:b]), np.asarray(d_got[a:b])
            assert np.array_equal(dr[oref].view(np.uint8), dg[ogot].view(np.uint8)), f"row {r}: distances differ"
            if sorted_got and b - a > 1:
                assert np.all((dg[1:] > dg[:-1]) | ((dg[1:] == dg[:-1]) & (ig[1:] > ig[:-1]))), \
                    f"row {r}: not sorted by (distance, index)"


def preorder(tree_dump):
    """Node ids of a GPU tree dump in preorder (node, left, right)."""
    child = tree_dump["child"]
    order = []
    stack = [0]
    while stack:
        v = stack.pop()
        order.append(v)
        if child[v] >= 0:
            stack.append(child[v] + 1)
            stack.append(child[v])
    return np.array(order, dtype=np.int64)


def assert_tree_equal(oracle_dump, gpu_dump):
    """Same tree node for node (preorder): leaf flag, axis, split value,
    cell bounds and the set of original indices in each leaf."""
    order = preorder(gpu_dump)
    on = len(oracle_dump["axis"])
    assert len(order) == on, f"node counts differ: oracle {on} gpu {len(order)}"
    g_leaf = gpu_dump["axis"][order] < 0
    assert np.array_equal(oracle_dump["is_leaf"].astype(bool), g_leaf), "leaf flags differ"
    inner = ~g_leaf
    assert np.array_equal(oracle_dump["axis"][inner], gpu_dump["axis"][order][inner]), "axes differ"
    if inner.any():  # a single-leaf tree has no split
        assert np.array_equal(oracle_dump["split"][inner].view(np.uint8).reshape(inner.sum(), -1),
                              gpu_dump["split"][order][inner].view(np.uint8).reshape(inner.sum(), -1)), \
            "splits differ"
    assert np.array_equal(oracle_dump["minb"], gpu_dump["lo"][order]), "cell lower bounds differ"
    assert np.array_equal(oracle_dump["maxb"], gpu_dump["hi"][order]), "cell upper bounds differ"
    assert np.array_equal(oracle_dump["count"], gpu_dump["count"][order]), "point counts differ"
    for j in np.nonzero(g_leaf)[0]:
        v = order[j]
        s, c = gpu_dump["start"][v], gpu_dump["count"][v]
        so, co = oracle_dump["start"][j], oracle_dump["count"][j]
        assert set(gpu_dump["idx"][s:s + c].tolist()) == set(oracle_dump["lidx"][so:so + co].tolist()), \
            f"leaf {j}: point sets differ"
