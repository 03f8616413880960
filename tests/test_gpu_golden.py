"""The reference's own outputs (tests/golden/g1-g6, written by the
reference compiled from src/cosmokdtree.f90) replayed on the GPU.

Every fixture input is built into a GPU tree and queried through the C ABI;
the results must equal the stored reference outputs:
  * kNN: distances bit for bit, index sets below the k-th distance, ties at
    the k-th distance as sets (cosmokdtree.f90:658), with every kNN kernel;
  * ball and box: the same (index, distance) sets per row, row counts
    included (strict child visits of box_search, cosmokdtree.f90:1296-1302,
    decide which points on split values are found);
  * tree dumps (g4): node for node, the inputs have no duplicate coordinates;
  * edge cases (g5): k > n sentinels (sqrt(HUGE(0.0)) in the tree precision,
    index -2: cosmokdtree.f90:517-518, kdtree.py:52), identical points,
    points on the periodic faces, boxes on split values, a radius equal to a
    distance (inclusive, cosmokdtree.f90:1126), empty results
    (:1049-1053, :1241-1243).
The drop-in functions of cosmokdtree_amd.kdtree are checked with float32
and float64 numpy input against the float64 oracle: the reference converts
every input to float64 (src/cosmokdtree/kdtree.py:31,43,59,83).
"""
import os

import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree
from tests.parity import assert_csr_equal, assert_knn_rows, assert_tree_equal

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KERNELS = [("auto", 0, 0), ("lane", 0, 0), ("wave", 0, 0), ("stack", 0, 0), ("stack", 0, 4)]
KIDS = ["auto", "lane", "wave", "stack", "stack_q4"]
SENTINEL = {np.float32: np.sqrt(np.float32(np.finfo(np.float32).max)),
            np.float64: np.sqrt(np.float64(np.finfo(np.float32).max))}


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def knn_gpu(tree, q, k, sorted=True):
    d, i = tree.knn(torch.from_numpy(np.ascontiguousarray(q)), k, sorted=sorted)
    return d.cpu().numpy(), i.cpu().numpy()


def ball_gpu(tree, q, r, sorted=False):
    off, d, i = tree.ball(torch.from_numpy(np.ascontiguousarray(q)), r, sorted=sorted)
    return off.cpu().numpy(), d.cpu().numpy(), i.cpu().numpy()


def box_gpu(tree, boxes):
    off, i = tree.box(torch.from_numpy(np.ascontiguousarray(boxes)))
    return off.cpu().numpy(), i.cpu().numpy()


@pytest.mark.parametrize("kern", KERNELS, ids=KIDS)
@pytest.mark.parametrize("name", ["g1_uniform_f64", "g3_uniform_f32", "g2_periodic_f64", "g3_periodic_f32"])
def test_reference_fixture(ckd, name, kern):
    g = load(name)
    pts, q = g["points"], g["queries"]
    L = g["boxsize"] if "boxsize" in g else None
    print(f"\n[golden {name} kernel {kern}] n={len(pts)} q={len(q)}", flush=True)
    tree = ckd.kdtree._build(pts, int(g["leaf"]), L)
    try:
        tree.knn_config(kern[0], kern[1], kern[2])
        for key in g.files:
            if key.startswith("knn") and key.endswith("_d"):
                k = int(key.split("_")[1][1:])
                s = key.split("_")[2] == "s"
                d, i = knn_gpu(tree, q, k, s)
                assert_knn_rows(pts, q, k, g[key], g[key[:-2] + "_i"], d, i, L=L, sorted_got=s)
        for j, r in enumerate(g["ball_radii"]):
            for s in (True, False):
                key = f"ball_r{j}_{'s' if s else 'u'}"
                off, d, i = ball_gpu(tree, q, float(r), s)
                assert_csr_equal(g[key + "_off"], g[key + "_d"], g[key + "_i"], off, d, i, sorted_got=s)
        off, i = box_gpu(tree, g["box_boxes"])
        assert_csr_equal(g["box_off"], None, g["box_i"], off, None, i)
    finally:
        tree.free()


def test_reference_tree_dumps(ckd):
    g = load("g4_trees")
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_points")})
    assert len(keys) == 36
    for key in keys:
        pts = g[key + "_points"]
        leaf = int(key.split("_l")[-1])
        L = np.full(3, 10.0) if key.startswith("p1") else None
        print(f"[tree {key}]", flush=True)
        tree = ckd.kdtree._build(pts, leaf, L)
        try:
            ref = {f: g[f"{key}_{f}"] for f in ("is_leaf", "axis", "split", "minb", "maxb", "start", "count",
                                                 "lidx")}
            assert_tree_equal(ref, tree.dump())
        finally:
            tree.free()


@pytest.mark.parametrize("kern", KERNELS, ids=KIDS)
def test_reference_edges_knn(ckd, kern):
    g = load("g5_edges")
    # k = n and k > n: the slots past n hold sqrt(HUGE(0.0)) and index -2
    pts, q = g["small_points"], g["small_queries"]
    tree = ckd.kdtree._build(pts, 4, None).knn_config(kern[0], kern[1], kern[2])
    try:
        for k in (10, 13):
            print(f"[edges small k={k} {kern}]", flush=True)
            d, i = knn_gpu(tree, q, k)
            assert_knn_rows(pts, q, k, g[f"small_k{k}_d"], g[f"small_k{k}_i"], d, i)
        d, i = knn_gpu(tree, q, 13)
        assert np.all(i[:, 10:] == -2) and np.all(d[:, 10:] == SENTINEL[np.float64])
    finally:
        tree.free()
    # the float32 tree's sentinel is sqrt(FLT_MAX) in float32
    tree = ckd.kdtree._build(pts.astype(np.float32), 4, None).knn_config(kern[0], kern[1], kern[2])
    try:
        d, i = knn_gpu(tree, q.astype(np.float32), 13)
        assert d.dtype == np.float32
        assert np.all(i[:, 10:] == -2) and np.all(d[:, 10:] == SENTINEL[np.float32])
    finally:
        tree.free()
    for part, k, leaf, L in (("same", 8, 4, None), ("dup", 16, 16, None), ("face", 12, 8, np.full(3, 4.0))):
        pts, q = g[f"{part}_points"], g[f"{part}_queries"]
        print(f"[edges {part} k={k} {kern}]", flush=True)
        tree = ckd.kdtree._build(pts, leaf, L).knn_config(kern[0], kern[1], kern[2])
        try:
            d, i = knn_gpu(tree, q, k)
            assert_knn_rows(pts, q, k, g[f"{part}_k{k}_d"], g[f"{part}_k{k}_i"], d, i, L=L)
        finally:
            tree.free()


@pytest.mark.parametrize("k", [70, 100])
def test_heap_kernel_k_above_n(ckd, k):
    """k > 64 runs the max-heap rows (max_heap_insert, cosmokdtree.f90:859-905)
    and the segmented sort; with k > n the empty slots follow the same rule."""
    g = load("g5_edges")
    pts, q = g["small_points"], g["small_queries"]
    orc = OracleTree(pts, 4)
    for dt in (np.float64, np.float32):
        tree = ckd.kdtree._build(pts.astype(dt), 4, None)
        try:
            for s in (True, False):
                d, i = knn_gpu(tree, q.astype(dt), k, s)
                if s:  # sorted: the empty slots come last
                    assert np.all(i[:, 10:] == -2) and np.all(d[:, 10:] == SENTINEL[dt])
                # unsorted rows keep the reference's heap layout: empty slots anywhere
                assert np.all((i == -2).sum(1) == k - 10) and np.all(d[i == -2] == SENTINEL[dt])
                assert np.all(d[i != -2] < SENTINEL[dt])
                if dt == np.float64:
                    dr, ir = orc.knn(q, k, s)
                    assert_knn_rows(pts, q, k, dr, ir, d, i, sorted_got=s)
                    assert np.array_equal(np.sort(ir, 1), np.sort(i, 1))
        finally:
            tree.free()
    # a larger tree: k > 64 with k <= n, against the oracle
    rng = np.random.default_rng(77)
    pts = rng.random((3000, 3))
    q = rng.random((300, 3))
    tree = ckd.kdtree._build(pts, 16, None)
    try:
        dr, ir = OracleTree(pts, 16).knn(q, k, True)
        d, i = knn_gpu(tree, q, k)
        assert_knn_rows(pts, q, k, dr, ir, d, i)
    finally:
        tree.free()


def test_reference_edges_ball_box(ckd):
    g = load("g5_edges")
    # identical points: one deep slide chain; every point at distance 0 or none
    pts, q = g["same_points"], g["same_queries"]
    tree = ckd.kdtree._build(pts, 4, None)
    try:
        off, d, i = ball_gpu(tree, q, 0.1, True)
        assert_csr_equal(g["same_ball_off"], g["same_ball_d"], g["same_ball_i"], off, d, i, sorted_got=True)
        dump = tree.dump()
        from tests.parity import preorder
        order = preorder(dump)
        # the slides peel one point per level whichever identical point they take
        assert np.array_equal(g["same_tree_is_leaf"].astype(bool), dump["axis"][order] < 0)
        assert np.array_equal(g["same_tree_count"], dump["count"][order])
        assert np.array_equal(g["same_tree_minb"], dump["lo"][order])
        assert np.array_equal(g["same_tree_maxb"], dump["hi"][order])
    finally:
        tree.free()
    # duplicates on one axis: boxes with edges on the grid values
    tree = ckd.kdtree._build(g["dup_points"], 16, None)
    try:
        off, i = box_gpu(tree, g["dup_boxes"])
        assert_csr_equal(g["dup_box_off"], None, g["dup_box_i"], off, None, i)
    finally:
        tree.free()
    # points exactly on the periodic faces
    tree = ckd.kdtree._build(g["face_points"], 8, np.full(3, 4.0))
    try:
        off, d, i = ball_gpu(tree, g["face_queries"], 0.3, True)
        assert_csr_equal(g["face_ball_off"], g["face_ball_d"], g["face_ball_i"], off, d, i, sorted_got=True)
    finally:
        tree.free()
    # box edges on split values (strict visits), a radius equal to a distance, empty results
    tree = ckd.kdtree._build(g["split_points"], 16, None)
    try:
        off, i = box_gpu(tree, g["split_boxes"])
        assert_csr_equal(g["split_box_off"], None, g["split_box_i"], off, None, i)
        q, radii = g["req_queries"], g["req_radii"]
        got_d, got_i, got_off = [], [], [0]
        for j in range(len(q)):
            o, d, i = ball_gpu(tree, q[j:j + 1], float(radii[j]), True)
            got_d.append(d)
            got_i.append(i)
            got_off.append(got_off[-1] + len(i))
        assert_csr_equal(g["req_off"], g["req_d"], g["req_i"], np.array(got_off), np.concatenate(got_d),
                         np.concatenate(got_i), sorted_got=True)
        off, d, i = ball_gpu(tree, np.array([[5.0, 5.0, 5.0]]), 0.5)
        assert np.array_equal(off, g["empty_ball_off"]) and len(i) == 0
        off, i = box_gpu(tree, np.array([[2.0, 3.0, 2.0, 3.0, 2.0, 3.0]]))
        assert np.array_equal(off, g["empty_box_off"]) and len(i) == 0
    finally:
        tree.free()


@pytest.mark.parametrize("d", [2, 6])
def test_reference_dims(ckd, d):
    g = load("g6_dims")
    pts, q = g[f"d{d}_points"], g[f"d{d}_queries"]
    tree = ckd.kdtree._build(pts, 16, None)
    try:
        dd, ii = knn_gpu(tree, q, 10)
        assert_knn_rows(pts, q, 10, g[f"d{d}_k10_d"], g[f"d{d}_k10_i"], dd, ii)
        off, bd, bi = ball_gpu(tree, q, float(g[f"d{d}_radius"]), True)
        assert_csr_equal(g[f"d{d}_ball_off"], g[f"d{d}_ball_d"], g[f"d{d}_ball_i"], off, bd, bi, sorted_got=True)
        off, bi = box_gpu(tree, g[f"d{d}_boxes"])
        assert_csr_equal(g[f"d{d}_box_off"], None, g[f"d{d}_box_i"], off, None, bi)
    finally:
        tree.free()


@pytest.mark.parametrize("in_dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", ["g1_uniform_f64", "g2_periodic_f64"])
def test_dropin_float64_semantics(ckd, name, in_dtype):
    """The five drop-in functions with float32 or float64 numpy input give
    the float64 results of the reference (src/cosmokdtree/kdtree.py:16-87):
    the input is converted to float64 first."""
    kd = ckd.kdtree
    g = load(name)
    pts_in = g["points"].astype(in_dtype)
    q_in = g["queries"].astype(in_dtype)
    L = g["boxsize"] if "boxsize" in g else None
    p64, q64 = pts_in.astype(np.float64), q_in.astype(np.float64)
    orc = OracleTree(p64, 16, L)
    tree = kd.build_kdtree(pts_in, boxsize=None if L is None else list(L))
    try:
        assert isinstance(tree, int)
        d, i = kd.knn_search(tree, q_in, 8)
        assert d.dtype == np.float64 and i.dtype == np.int64 and d.flags.f_contiguous
        dr, ir = orc.knn(q64, 8, True)
        assert_knn_rows(p64, q64, 8, dr, ir, d, i, L=L)
        d1, i1 = kd.knn_search(tree, q_in[0], np.int32(8))  # one 1-D target
        assert d1.shape == (1, 8) and np.array_equal(d1[0], d[0])
        r = float(g["ball_radii"][0])
        dl, il = kd.ball_search(tree, q_in, np.float32(r) if in_dtype == np.float32 else r)
        assert len(dl) == len(q_in) and all(x.dtype == np.float64 for x in dl)
        rr = float(np.float32(r)) if in_dtype == np.float32 else r
        off, dd, ii = orc.ball_csr(q64, rr)
        goff = np.concatenate([[0], np.cumsum([len(x) for x in il])])
        assert_csr_equal(off, dd, ii, goff, np.concatenate(dl), np.concatenate(il))
        box = g["box_boxes"][0].astype(in_dtype)
        got = kd.box_search(tree, box)
        assert got.dtype == np.int64
        assert sorted(got.tolist()) == sorted(orc.box(box.astype(np.float64)).tolist())
    finally:
        kd.deallocate_kdtree(tree)
    with pytest.raises(ValueError):
        kd.knn_search(tree, q_in, 8)
    # float32 trees only on request
    t32 = kd.build_kdtree(pts_in, boxsize=None if L is None else list(L), dtype=np.float32)
    try:
        d, i = kd.knn_search(t32, q_in, 8)
        assert d.dtype == np.float32
        p32, q32 = pts_in.astype(np.float32), q_in.astype(np.float32)
        dr, ir = OracleTree(p32, 16, L).knn(q32, 8, True)
        assert_knn_rows(p32, q32, 8, dr, ir, d, i, L=L)
    finally:
        kd.deallocate_kdtree(t32)
