"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(ckd_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("ckd_build", "ckd_free", "ckd_knn", "ckd_ball_count", "ckd_ball_fill", "ckd_box_count",
                 "ckd_box_fill", "ckd_last_error"):
        assert must in names


def test_library_exports_all_declared_symbols():
    import torch  # noqa: F401  (same HIP runtime as in use)
    path = os.path.join(ROOT, "cosmokdtree_amd", "libcosmokdtree_amd.so")
    if not os.path.exists(path):
        pytest.fail("libcosmokdtree_amd.so is not built (python -m cosmokdtree_amd.build)")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    lib.ckd_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.ckd_version()


def test_package_imports_and_validates_arguments():
    import numpy as np

    from cosmokdtree_amd import kdtree
    with pytest.raises(AssertionError):
        kdtree.build_kdtree(np.zeros((4, 3)), leaf=0)
    with pytest.raises(AssertionError):
        kdtree.build_kdtree(np.zeros((4, 3)), boxsize=[1.0, 1.0])
    with pytest.raises(ValueError):
        kdtree.knn_search(12345, np.zeros(3), 3)
    with pytest.raises(AssertionError):
        kdtree.knn_search(12345, np.zeros(3), 3.0)
    with pytest.raises(AssertionError):
        kdtree.ball_search(12345, np.zeros(3), -1.0)


def test_tree_load_rejects_other_files(tmp_path):
    """A file that is not a saved tree is refused before any device call."""
    import ctypes
    path = os.path.join(ROOT, "cosmokdtree_amd", "libcosmokdtree_amd.so")
    lib = ctypes.CDLL(path)
    lib.ckd_tree_load.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.ckd_last_error.restype = ctypes.c_char_p
    bad = tmp_path / "not_a_tree.bin"
    bad.write_bytes(b"definitely not a tree" * 10)
    out = ctypes.c_void_p()
    assert lib.ckd_tree_load(str(bad).encode(), None, ctypes.byref(out)) == -1
    assert b"not a saved" in lib.ckd_last_error()
    assert lib.ckd_tree_load(str(tmp_path / "missing").encode(), None, ctypes.byref(out)) == -1


def test_batch_plan_limits():
    """The batched engine's 32-bit row limits are checked before any device
    work: forests or query batches of 2^31 - 1 rows and more are refused."""
    from cosmokdtree_amd.batch import check_plan_limits
    check_plan_limits(1000, [0, 10, 20])
    with pytest.raises(ValueError):
        check_plan_limits(2 ** 31 - 1, [0, 10])
    with pytest.raises(ValueError):
        check_plan_limits(1000, [0, 2 ** 30, 2 ** 31])
    with pytest.raises(ValueError):
        check_plan_limits(1000, [0, 10, 5])


def _tree_file(path, nodes, depth=1, nleaves=None, header=None):
    """A saved-tree file: header (ckd_persist.hip PersistHeader) + a float32
    node array; the rest of the arrays are never reached by these cases."""
    import struct
    h = dict(n=4, nnodes=len(nodes), nleaves=nleaves if nleaves is not None else sum(1 for nd in nodes if nd[1] < 0),
             nsets=1, nwide=-1, ndim=3, prec=4, leafsize=2, periodic=0, idx_bits=32, depth=depth)
    h.update(header or {})
    raw = b"CKDTREE1" + struct.pack("<5q6i8d", h["n"], h["nnodes"], h["nleaves"], h["nsets"], h["nwide"],
                                    h["ndim"], h["prec"], h["leafsize"], h["periodic"], h["idx_bits"], h["depth"],
                                    *h.get("L", [1.0] * 8))
    for split, axis, child, parent in nodes:
        raw += struct.pack("<f3i", split, axis, child, parent)
    path.write_bytes(raw)
    return path


def test_tree_load_rejects_corrupt_trees(tmp_path):
    """Headers and node links that a kernel would follow out of bounds are
    refused on the host, before the arrays reach the device."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "cosmokdtree_amd", "libcosmokdtree_amd.so"))
    lib.ckd_tree_load.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.ckd_last_error.restype = ctypes.c_char_p
    good = [(0.5, 0, 1, -1), (0.0, -1, -1, 0), (0.0, -1, -1, 0)]
    cases = {
        "negative depth": (dict(nodes=good, depth=-1), b"inconsistent tree header"),
        "leaf size 0": (dict(nodes=good, header=dict(leafsize=0)), b"inconsistent tree header"),
        "periodic without a box": (dict(nodes=good, header=dict(periodic=1, L=[0.0] * 8)), b"inconsistent"),
        "more wide records than nodes": (dict(nodes=good, header=dict(nwide=3)), b"inconsistent"),
        "child out of range": (dict(nodes=[(0.5, 0, 2, -1), (0.0, -1, -1, 0), (0.0, -1, -1, 0)]),
                               b"child index out of range"),
        "axis beyond ndim": (dict(nodes=[(0.5, 5, 1, -1), (0.0, -1, -1, 0), (0.0, -1, -1, 0)]), b"split axis"),
        "wrong parent": (dict(nodes=[(0.5, 0, 1, -1), (0.0, -1, -1, 0), (0.0, -1, -1, 2)]), b"do not form a tree"),
        "leaf count": (dict(nodes=good, nleaves=1), b"leaf count"),
        "truncated": (dict(nodes=good[:2], header=dict(nnodes=3)), b"truncated"),
    }
    for name, (kw, msg) in cases.items():
        p = _tree_file(tmp_path / (name.replace(" ", "_") + ".bin"), **kw)
        out = ctypes.c_void_p()
        assert lib.ckd_tree_load(str(p).encode(), None, ctypes.byref(out)) == -1, name
        assert msg in lib.ckd_last_error(), (name, lib.ckd_last_error())
        assert not out.value
