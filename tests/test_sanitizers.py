"""The checkers and the library's host code under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only).

The oracle is rebuilt with -fsanitize=address,undefined (oracle/Makefile,
target `sanitize`) and driven through every routine -- builds of all kinds
(periodic or not, float32/float64, leaf 1 and 16, one point, identical
points), kNN with k below, at and above n (sorted and unsorted), ball, box
and wrapped-box search, tree dumps -- in a child process with the ASan
runtime preloaded; any report fails the test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

EXERCISE = r'''
import numpy as np
from oracle.oracle import OracleTree
rng = np.random.default_rng(3)
for dt in (np.float32, np.float64):
    for periodic in (False, True):
        for n, leaf in ((1, 16), (7, 1), (500, 1), (3000, 16)):
            L = np.full(3, 10.0) if periodic else None
            pts = ((rng.random((n, 3)) - 0.5) * 10).astype(dt)
            t = OracleTree(pts, leaf, L)
            q = ((rng.random((40, 3)) - 0.5) * 10).astype(dt)
            for k in (1, min(n, 8), n + 3, 70):
                for s in (True, False):
                    t.knn(q, k, s)
            t.ball_csr(q, 0.7, True)
            t.ball_csr(q, 0.7, False)
            b = np.concatenate([q - 0.9, q + 0.9], 1)[:, [0, 3, 1, 4, 2, 5]]
            t.box_csr(b)
            if periodic:
                b[:, 0] -= 4.0
                t.box_csr(b, wrap=True)
            t.dump()
            t.close()
same = np.full((60, 3), 0.25)
t = OracleTree(same, 4)
t.knn(same[:3], 10, True)
t.ball_csr(same[:3], 0.1, True)
t.dump()
print("exercised")
'''


def _asan_lib():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if p.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_under_asan_and_ubsan(tmp_path):
    asan = _asan_lib()
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    out = str(tmp_path)
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize", f"OUT={out}"], check=True,
                   capture_output=True)
    env = dict(os.environ, LD_PRELOAD=asan, CKD_ORACLE_SO=os.path.join(out, "libckd_oracle.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               OMP_NUM_THREADS="2")
    p = subprocess.run([sys.executable, "-c", EXERCISE], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=300)
    report = p.stderr[-4000:]
    assert p.returncode == 0 and "exercised" in p.stdout, report
    assert "AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr, report


LIB_EXERCISE = r'''
import ctypes as C, os, struct, tempfile
from cosmokdtree_amd._lib import lib, LIB_PATH
assert LIB_PATH.endswith("_debug.so"), LIB_PATH
vp, i64 = C.c_void_p, C.c_int64
lib.ckd_last_error.restype = C.c_char_p
out = vp()
# argument checks of the entry points (rejected before any device call)
assert lib.ckd_build(None, 10, 3, 4, 16, None, 0, None, C.byref(out)) == -1
assert lib.ckd_build(vp(1), 10, 9, 4, 16, None, 0, None, C.byref(out)) == -1
assert lib.ckd_build(vp(1), 10, 3, 2, 16, None, 0, None, C.byref(out)) == -1
assert lib.ckd_build(vp(1), 10, 3, 4, 0, None, 0, None, C.byref(out)) == -1
assert lib.ckd_build(vp(1), 1 << 31, 3, 4, 16, None, 32, None, C.byref(out)) == -5
bad = (C.c_double * 3)(1.0, -1.0, 1.0)
assert lib.ckd_build(vp(1), 10, 3, 4, 16, bad, 0, None, C.byref(out)) == -1
offs = (C.c_int64 * 3)(0, 5, 4)
assert lib.ckd_build_forest(vp(1), offs, 2, 3, 4, 16, None, 0, None, C.byref(out)) == -1
assert lib.ckd_knn(None, None, 1, 4, 1, None, None, None, None, None) != 0
assert lib.ckd_ball_count(None, None, 1, 1.0, None, None, None) != 0
assert lib.ckd_box_count(None, None, 1, None, None, None) != 0
assert lib.ckd_box_count_periodic(None, None, 1, None, None, None) != 0
assert lib.ckd_knn_noalloc(None, None, 1, 4, 1, None, None, None, None, None) != 0
assert lib.ckd_batch_plan(None, None, None, None, 1, 256, 0, 0, C.byref(out), None) != 0
assert lib.ckd_free(None) == 0
assert len(lib.ckd_last_error()) > 0
# the tree loader on files that are not, or not quite, saved trees
def tree_file(path, nodes, nleaves, nnodes=None, depth=1, leafsize=2, periodic=0, L=1.0):
    raw = b"CKDTREE1" + struct.pack("<5q6i8d", 4, nnodes or len(nodes), nleaves, 1, -1, 3, 4, leafsize, periodic,
                                    32, depth, *([L] * 8))
    for sp, ax, ch, pa in nodes:
        raw += struct.pack("<f3i", sp, ax, ch, pa)
    open(path, "wb").write(raw)
good = [(0.5, 0, 1, -1), (0.0, -1, -1, 0), (0.0, -1, -1, 0)]
d = tempfile.mkdtemp()
cases = [dict(nodes=good, nleaves=2, depth=-1), dict(nodes=good, nleaves=2, leafsize=0),
         dict(nodes=good, nleaves=2, periodic=1, L=0.0), dict(nodes=[(0.5, 0, 2, -1)] + good[1:], nleaves=2),
         dict(nodes=[(0.5, 7, 1, -1)] + good[1:], nleaves=2), dict(nodes=good[:2], nleaves=2, nnodes=3),
         dict(nodes=good, nleaves=1), dict(nodes=[(0.5, 0, 1, -1), (0.0, -1, -1, 2), (0.0, -1, -1, 0)], nleaves=2)]
for j, kw in enumerate(cases):
    p = os.path.join(d, f"t{j}.bin")
    tree_file(p, **kw)
    assert lib.ckd_tree_load(p.encode(), None, C.byref(out)) == -1, j
open(os.path.join(d, "junk"), "wb").write(os.urandom(4096))
assert lib.ckd_tree_load(os.path.join(d, "junk").encode(), None, C.byref(out)) == -1
assert lib.ckd_tree_load(b"/nonexistent/tree", None, C.byref(out)) == -1
print("exercised")
'''


def test_library_host_code_under_asan_and_ubsan():
    """The debug build of the library (python -m cosmokdtree_amd.build
    --debug, also made by build()): its host code -- argument checks of the
    entry points and the tree loader on corrupt files -- under ASan and
    UBSan, in a child process with the clang ASan runtime preloaded."""
    from cosmokdtree_amd import build as B
    if not os.path.exists(B.DEBUG_LIB):
        pytest.fail(f"{B.DEBUG_LIB} is not built (python -m cosmokdtree_amd.build --debug)")
    env = dict(os.environ, LD_PRELOAD=B.asan_runtime(), CKD_LIB=B.DEBUG_LIB,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([sys.executable, "-c", LIB_EXERCISE], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=300)
    report = p.stderr[-4000:]
    assert p.returncode == 0 and "exercised" in p.stdout, report
    assert "AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr, report
