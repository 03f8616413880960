"""The Fortran module cosmokdtree_gpu (cosmokdtree_amd/fortran) against the
oracle, through the check program tests/fortran/ckd_fortran_check.F90 built
for three compile-time variants (periodic float32 int32 indices, the
reference's default kinds; non-periodic float64 int64 indices, the Python
build's kinds; a 2-D float64 tree).

The program calls the module as the reference's Fortran demos call module
cosmokdtree (build_kdtree, knn_search, ball_search, box_search,
deallocate_query, deallocate_kdtree) and the module's batched forms, and
writes the results; this test compares them with the oracle: kNN to the
kNN parity bar, ball and box as sets with bit-identical distances, the
single-target calls against the batched ones, 1-based indices.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleTree
from tests.parity import assert_csr_equal, assert_knn_rows

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "fortran", "bin")

# name: periodic, ndim, dtype
VARIANTS = {
    "p1_d3_f32_i4": (True, 3, np.float32),
    "p0_d3_f64_i8": (False, 3, np.float64),
    "p0_d2_f64_i4": (False, 2, np.float64),
}


def _read(path, dtype):
    return np.fromfile(path, dtype=dtype)


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_fortran_module(ckd, tmp_path, variant):
    exe = os.path.join(BIN, f"ckd_fortran_check_{variant}")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (make -C tests/fortran)")
    periodic, D, dt = VARIANTS[variant]
    rng = np.random.default_rng(31)
    n, nq, k, nsingle, leaf = 30000, 2000, 16, 40, 16
    if periodic:
        L = np.full(D, 10.0)
        pts = ((rng.random((n, D)) - 0.5) * 10.0).astype(dt)
        q = np.clip(pts[rng.integers(0, n, nq)] + (rng.random((nq, D)) - 0.5) * 0.4, -5.0, 5.0).astype(dt)
        radius, side = 0.3, 0.6
    else:
        L = None
        pts = rng.random((n, D)).astype(dt)
        q = rng.random((nq, D)).astype(dt)
        radius, side = 0.02, 0.04
    c = rng.random((60, D)) * (10.0 if periodic else 1.0) - (5.0 if periodic else 0.0)
    boxes = np.empty((60, 2 * D), dtype=dt)
    boxes[:, 0::2] = c - side / 2
    boxes[:, 1::2] = c + side / 2
    with open(tmp_path / "meta.txt", "w") as f:
        f.write(f"{n} {nq} {k} {radius!r} {len(boxes)} {nsingle} {leaf}\n")
        if periodic:
            f.write(" ".join(repr(float(x)) for x in L) + "\n")
    np.ascontiguousarray(pts.T).tofile(tmp_path / "pts.bin")  # Fortran order (n, ndim)
    q.tofile(tmp_path / "q.bin")
    boxes.tofile(tmp_path / "box.bin")
    print(f"\n[fortran {variant}] n={n} nq={nq} k={k}", flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "done" in r.stdout, r.stdout + r.stderr

    orc = OracleTree(pts, leaf, L)
    try:
        # kNN: sorted rows, 1-based indices
        d = _read(tmp_path / "knn_d.bin", dt).reshape(nq, k)
        i = _read(tmp_path / "knn_i.bin", np.int64).reshape(nq, k)
        assert i.min() >= 1 and i.max() <= n
        d_ref, i_ref = orc.knn(q, k, True)
        assert_knn_rows(pts, q, k, d_ref, i_ref, d, i - 1, L=L)
        d1 = _read(tmp_path / "knn1_d.bin", dt).reshape(nsingle, k)
        i1 = _read(tmp_path / "knn1_i.bin", np.int64).reshape(nsingle, k)
        np.testing.assert_array_equal(d1, d[:nsingle])
        np.testing.assert_array_equal(i1, i[:nsingle])
        # ball: batched (sorted) against the oracle; single calls (unsorted)
        # against the batched rows as sets
        off = _read(tmp_path / "ball_off.bin", np.int64)
        bd = _read(tmp_path / "ball_d.bin", dt)
        bi = _read(tmp_path / "ball_i.bin", np.int64)
        o_ref, bd_ref, bi_ref = orc.ball_csr(q, radius, True)
        assert_csr_equal(o_ref, bd_ref, bi_ref, off, bd, bi - 1, sorted_got=True)
        cnt1 = _read(tmp_path / "ball1_n.bin", np.int64)
        off1 = np.concatenate([[0], np.cumsum(cnt1)])
        assert_csr_equal(off[:nsingle + 1], bd, bi, off1, _read(tmp_path / "ball1_d.bin", dt),
                         _read(tmp_path / "ball1_i.bin", np.int64))
        # box
        off = _read(tmp_path / "box_off.bin", np.int64)
        xi = _read(tmp_path / "box_i.bin", np.int64)
        ob, ib = orc.box_csr(boxes)
        assert off[-1] > 0
        assert_csr_equal(ob, None, ib, off, None, xi - 1)
        cnt1 = _read(tmp_path / "box1_n.bin", np.int64)
        off1 = np.concatenate([[0], np.cumsum(cnt1)])
        m = len(cnt1)
        assert_csr_equal(off[:m + 1], None, xi, off1, None, _read(tmp_path / "box1_i.bin", np.int64))
    finally:
        orc.close()
