"""cosmokdtree_amd -- MI355X (gfx950) k-d tree build and query path.

Drop-in for cosmokdtree's Python API (src/cosmokdtree/kdtree.py):
``from cosmokdtree_amd import kdtree`` then the same calls.

The native library is loaded on first use of the API (not at package
import), so ``python -m cosmokdtree_amd.build`` works before it exists.
"""
__version__ = "0.1.0"

_API = ("KDTree", "ball_search", "box_search", "box_search_batch", "build_kdtree", "deallocate_kdtree",
        "knn_search")


_MODULES = ("kdtree", "learned", "batch", "synthetic", "build", "autotune", "largebox", "pycosmokdtree")


def __getattr__(name):
    import importlib
    if name in _API:
        return getattr(importlib.import_module(".kdtree", __name__), name)
    if name in _MODULES:
        return importlib.import_module("." + name, __name__)
    raise AttributeError(f"module 'cosmokdtree_amd' has no attribute {name!r}")
