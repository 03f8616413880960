"""Pin the C oracle to the reference's golden vectors (tests/golden/*.npz).

The fixtures were written by the reference Fortran module itself
(tests/golden/make_golden.py).  The oracle restates the same algorithm
step for step, so it must reproduce them exactly: same tree (node for node,
leaf point order included), same kNN rows including the unsorted heap
layout, same ball and box rows including traversal order.
"""
import os

import numpy as np
import pytest

from oracle.oracle import OracleTree

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name", ["g1_uniform_f64", "g3_uniform_f32", "g2_periodic_f64", "g3_periodic_f32"])
def test_oracle_reproduces_reference(name):
    g = load(name)
    pts, q = g["points"], g["queries"]
    L = g["boxsize"] if "boxsize" in g else None
    t = OracleTree(pts, int(g["leaf"]), L)
    for key in g.files:
        if key.startswith("knn") and key.endswith("_d"):
            k = int(key.split("_")[1][1:])
            s = key.split("_")[2] == "s"
            d, i = t.knn(q, k, s)
            assert same(d, g[key]), key
            assert same(i, g[key[:-2] + "_i"]), key
    for j, r in enumerate(g["ball_radii"]):
        for s in (True, False):
            key = f"ball_r{j}_{'s' if s else 'u'}"
            off, d, i = t.ball_csr(q, float(r), s)
            assert same(off, g[key + "_off"]) and same(d, g[key + "_d"]) and same(i, g[key + "_i"]), key
    off, i = t.box_csr(g["box_boxes"])
    assert same(off, g["box_off"]) and same(i, g["box_i"])


def test_oracle_tree_dumps():
    g = load("g4_trees")
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_points")})
    assert len(keys) == 36
    for key in keys:
        pts = g[key + "_points"]
        periodic = key.startswith("p1")
        leaf = int(key.split("_l")[-1])
        t = OracleTree(pts, leaf, np.full(3, 10.0) if periodic else None)
        d = t.dump()
        for f in ("is_leaf", "axis", "split", "minb", "maxb", "left", "right", "start", "count", "lidx"):
            if f in ("axis", "split"):  # leaves carry no split in the reference
                inner = d["is_leaf"] == 0
                assert same(d[f][inner], g[f"{key}_{f}"][inner]), (key, f)
            else:
                assert same(d[f], g[f"{key}_{f}"]), (key, f)


def test_oracle_edges():
    g = load("g5_edges")
    t = OracleTree(g["small_points"], 4)
    for k in (10, 13):
        d, i = t.knn(g["small_queries"], k, True)
        assert same(d, g[f"small_k{k}_d"]) and same(i, g[f"small_k{k}_i"])
    d, i = t.knn(g["small_queries"], 13, True)
    assert np.all(i[:, 10:] == -2)  # k > n: empty slots are -2 after idx -= 1
    assert np.all(d[:, 10:] == np.sqrt(np.float64(np.finfo(np.float32).max)))
    t = OracleTree(g["same_points"], 4)
    d, i = t.knn(g["same_queries"], 8, True)
    assert same(d, g["same_k8_d"]) and same(i, g["same_k8_i"])
    off, dd, ii = t.ball_csr(g["same_queries"], 0.1, True)
    assert same(off, g["same_ball_off"]) and same(dd, g["same_ball_d"]) and same(ii, g["same_ball_i"])
    dump = t.dump()
    for f in ("is_leaf", "minb", "maxb", "count", "lidx"):
        assert same(dump[f], g["same_tree_" + f]), f
    t = OracleTree(g["dup_points"], 16)
    d, i = t.knn(g["dup_queries"], 16, True)
    assert same(d, g["dup_k16_d"]) and same(i, g["dup_k16_i"])
    off, ii = t.box_csr(g["dup_boxes"])
    assert same(off, g["dup_box_off"]) and same(ii, g["dup_box_i"])
    t = OracleTree(g["face_points"], 8, np.full(3, 4.0))
    d, i = t.knn(g["face_queries"], 12, True)
    assert same(d, g["face_k12_d"]) and same(i, g["face_k12_i"])
    off, dd, ii = t.ball_csr(g["face_queries"], 0.3, True)
    assert same(off, g["face_ball_off"]) and same(dd, g["face_ball_d"]) and same(ii, g["face_ball_i"])
    t = OracleTree(g["split_points"], 16)
    off, ii = t.box_csr(g["split_boxes"])
    assert same(off, g["split_box_off"]) and same(ii, g["split_box_i"])
    q, radii = g["req_queries"], g["req_radii"]
    got_i, got_d = [], []
    for j in range(len(q)):
        o, dd, ii = t.ball_csr(q[j:j + 1], float(radii[j]), True)
        got_i.append(ii)
        got_d.append(dd)
    assert same(np.concatenate(got_i), g["req_i"]) and same(np.concatenate(got_d), g["req_d"])
    off, _, _ = t.ball_csr(np.array([[5.0, 5.0, 5.0]]), 0.5)
    assert same(off, g["empty_ball_off"]) and off[-1] == 0


@pytest.mark.parametrize("d", [2, 6])
def test_oracle_dims(d):
    g = load("g6_dims")
    t = OracleTree(g[f"d{d}_points"], 16)
    q = g[f"d{d}_queries"]
    dd, ii = t.knn(q, 10, True)
    assert same(dd, g[f"d{d}_k10_d"]) and same(ii, g[f"d{d}_k10_i"])
    off, bd, bi = t.ball_csr(q, float(g[f"d{d}_radius"]), True)
    assert same(off, g[f"d{d}_ball_off"]) and same(bd, g[f"d{d}_ball_d"]) and same(bi, g[f"d{d}_ball_i"])
    off, bi = t.box_csr(g[f"d{d}_boxes"])
    assert same(off, g[f"d{d}_box_off"]) and same(bi, g[f"d{d}_box_i"])
