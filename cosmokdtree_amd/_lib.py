"""ctypes binding of ``libcosmokdtree_amd.so`` (include/cosmokdtree_amd.h).

The library is loaded after ``torch`` so that both share one HIP runtime
(torch's bundled ``libamdhip64.so.7`` satisfies the library's dependency by
soname).  There is no CPU fallback: if the library is missing or cannot be
loaded, importing this module raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (loads the HIP runtime first)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CKD_LIB") or os.path.join(HERE, "libcosmokdtree_amd.so")

CKD_OK = 0
CKD_ERR_ARG = -1
CKD_ERR_RANGE = -2
CKD_ERR_HIP = -3
CKD_ERR_MEMORY = -4
CKD_ERR_LIMIT = -5
CKD_EMPTY_SLOT = -2
CKD_MAX_DIM = 8

EXPORTS = (
    "ckd_build", "ckd_build_forest", "ckd_free", "ckd_get_info", "ckd_knn", "ckd_ball_count", "ckd_ball_fill",
    "ckd_box_count", "ckd_box_fill", "ckd_tree_dump", "ckd_last_error", "ckd_version", "ckd_knn_config",
    "ckd_knn_timing", "ckd_knn_learned", "ckd_mlp_table", "ckd_tree_save", "ckd_tree_load", "ckd_sph_density",
    "ckd_release_memory", "ckd_ball_search",
)
KNN_KERNELS = {"auto": -1, "lane": 0, "wave": 1, "stack": 2}


class TreeInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("ndim", C.c_int32), ("prec", C.c_int32), ("leafsize", C.c_int32),
        ("periodic", C.c_int32), ("idx_bits", C.c_int32), ("depth", C.c_int32), ("nnodes", C.c_int64),
        ("nleaves", C.c_int64), ("boxsize", C.c_double * CKD_MAX_DIM), ("build_ms", C.c_double),
        ("nsets", C.c_int64),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m cosmokdtree_amd.build` (there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int32, C.c_double
    lib.ckd_build.argtypes = [vp, i64, i32, i32, i32, vp, i32, vp, C.POINTER(vp)]
    lib.ckd_build_forest.argtypes = [vp, vp, i64, i32, i32, i32, vp, i32, vp, C.POINTER(vp)]
    lib.ckd_free.argtypes = [vp]
    lib.ckd_get_info.argtypes = [vp, C.POINTER(TreeInfo)]
    lib.ckd_knn.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp, vp, vp]
    lib.ckd_knn_config.argtypes = [vp, i32, i32, i32]
    lib.ckd_knn_noalloc.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp, vp, vp]
    lib.ckd_knn_noalloc_scratch_bytes.argtypes = [vp, i64, i32]
    lib.ckd_knn_noalloc_scratch_bytes.restype = C.c_size_t
    lib.ckd_sph_density.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp]
    lib.ckd_knn_timing.argtypes = [vp, i32, C.POINTER(C.c_double), C.POINTER(i64)]
    lib.ckd_tree_save.argtypes = [vp, C.c_char_p]
    lib.ckd_tree_load.argtypes = [C.c_char_p, vp, C.POINTER(vp)]
    lib.ckd_ball_count.argtypes = [vp, vp, i64, dbl, vp, C.POINTER(i64), vp]
    lib.ckd_ball_fill.argtypes = [vp, vp, i64, dbl, i32, vp, vp, vp, vp]
    lib.ckd_ball_search.argtypes = [vp, vp, i64, dbl, i32, vp, i64, vp, vp, C.POINTER(i64), vp]
    lib.ckd_box_count.argtypes = [vp, vp, i64, vp, C.POINTER(i64), vp]
    lib.ckd_box_fill.argtypes = [vp, vp, i64, vp, vp, vp]
    lib.ckd_box_count_periodic.argtypes = [vp, vp, i64, vp, C.POINTER(i64), vp]
    lib.ckd_box_fill_periodic.argtypes = [vp, vp, i64, vp, vp, vp]
    lib.ckd_tree_dump.argtypes = [vp] * 11
    lib.ckd_release_memory.argtypes = []
    lib.ckd_last_error.restype = C.c_char_p
    lib.ckd_version.restype = C.c_char_p
    for name in EXPORTS:
        if name not in ("ckd_last_error", "ckd_version"):
            getattr(lib, name).restype = C.c_int
    return lib


lib = _load()


def check(rc: int, what: str = "") -> None:
    """Raise the Python exception matching a library status code."""
    if rc == CKD_OK:
        return
    msg = lib.ckd_last_error().decode(errors="replace") or f"error {rc}"
    if what:
        msg = f"{what}: {msg}"
    if rc in (CKD_ERR_ARG, CKD_ERR_RANGE):
        raise ValueError(msg)
    if rc == CKD_ERR_MEMORY:
        raise MemoryError(msg)
    raise RuntimeError(msg)
