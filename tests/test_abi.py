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
