"""GPU parity against the CPU oracle (oracle/ckd_oracle.c) on seeded inputs.

Each case builds the same points on the GPU (through the C ABI via the
drop-in package) and in the oracle, then compares:
  * tree structure node for node (duplicate-free inputs),
  * kNN rows (distances bit-exact, ties at the k-th distance as sets),
  * ball and box rows (same index sets, distances bit-exact).
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree
from tests.parity import assert_csr_equal, assert_knn_rows, assert_tree_equal

pytestmark = pytest.mark.gpu

CASES = [
    # name, n, d, dtype, periodic, leaf
    ("u3_f64", 20000, 3, np.float64, False, 16),
    ("u3_f32", 20000, 3, np.float32, False, 16),
    ("p3_f64", 20000, 3, np.float64, True, 16),
    ("p3_f32", 20000, 3, np.float32, True, 16),
    ("u2_f64_leaf1", 3000, 2, np.float64, False, 1),
    ("u6_f64", 5000, 6, np.float64, False, 16),
    ("clu3_f32", 30000, 3, np.float32, True, 16),
]


def make_points(name, n, d, dtype, periodic, seed=7):
    rng = np.random.default_rng(seed)
    if name.startswith("clu"):
        from cosmokdtree_amd.synthetic import clustered_periodic
        return clustered_periodic(n, 100.0, 0.5, seed=seed, dtype=dtype), np.full(3, 100.0)
    if periodic:
        L = np.full(d, 100.0)
        return ((rng.random((n, d)) - 0.5) * 100.0).astype(dtype), L
    return rng.random((n, d)).astype(dtype), None


def queries_for(pts, L, nq=1500, seed=11):
    rng = np.random.default_rng(seed)
    q = pts[rng.integers(0, len(pts), nq)].astype(np.float64)
    scale = 0.5 if L is not None else 0.005
    q = q + (rng.random(q.shape) - 0.5) * scale
    if L is not None:
        q = np.clip(q, -L / 2, L / 2)
    return q.astype(pts.dtype)


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request, ckd):
    name, n, d, dtype, periodic, leaf = request.param
    pts, L = make_points(name, n, d, dtype, periodic)
    print(f"\n[case {name}] n={n} d={d} {np.dtype(dtype).name} periodic={periodic} leaf={leaf}", flush=True)
    tree = ckd.kdtree._build(pts, leaf, L)
    orc = OracleTree(pts, leaf, L)
    q = queries_for(pts, L)
    yield dict(name=name, pts=pts, L=L, tree=tree, orc=orc, q=q, leaf=leaf)
    tree.free()
    orc.close()


def test_tree_structure(case):
    assert_tree_equal(case["orc"].dump(), case["tree"].dump())


@pytest.mark.parametrize("k", [1, 8, 32, 100])
def test_knn(case, k):
    q = case["q"]
    d_ref, i_ref = case["orc"].knn(q, k, True)
    d_got, i_got = case["tree"].knn(torch.from_numpy(q), k, sorted=True)
    assert_knn_rows(case["pts"], q, k, d_ref, i_ref, d_got.cpu().numpy(), i_got.cpu().numpy(), L=case["L"])


def test_knn_unsorted(case):
    q = case["q"]
    d_ref, i_ref = case["orc"].knn(q, 16, False)
    d_got, i_got = case["tree"].knn(torch.from_numpy(q), 16, sorted=False)
    assert_knn_rows(case["pts"], q, 16, d_ref, i_ref, d_got.cpu().numpy(), i_got.cpu().numpy(), L=case["L"],
                    sorted_got=False)


def _radius(case):
    if case["L"] is not None:
        return 2.0
    return 0.3 if case["pts"].shape[1] == 6 else 0.02


@pytest.mark.parametrize("kernel", ["stack", "lane"])
@pytest.mark.parametrize("sorted_", [False, True])
def test_ball(case, sorted_, kernel, monkeypatch):
    monkeypatch.setenv("CKD_BALL_KERNEL", kernel)  # walk per wave / one walk per query
    q = case["q"]
    r = _radius(case)
    off_r, d_r, i_r = case["orc"].ball_csr(q, r, sorted_)
    off, d, i = case["tree"].ball(torch.from_numpy(q), r, sorted=sorted_)
    assert_csr_equal(off_r, d_r, i_r, off.cpu().numpy(), d.cpu().numpy(), i.cpu().numpy(), sorted_got=sorted_)


def test_box(case):
    q = case["q"].astype(np.float64)
    d = q.shape[1]
    h = 1.5 if case["L"] is not None else 0.02
    boxes = np.empty((len(q), 2 * d))
    boxes[:, 0::2] = q - h
    boxes[:, 1::2] = q + h
    boxes = boxes.astype(case["pts"].dtype)
    off_r, i_r = case["orc"].box_csr(boxes)
    off, i = case["tree"].box(torch.from_numpy(boxes))
    assert_csr_equal(off_r, None, i_r, off.cpu().numpy(), None, i.cpu().numpy())
