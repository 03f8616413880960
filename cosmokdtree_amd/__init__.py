"""cosmokdtree_amd -- MI355X (gfx950) k-d tree build and query path.

Drop-in for cosmokdtree's Python API (src/cosmokdtree/kdtree.py):
``from cosmokdtree_amd import kdtree`` then the same calls.
"""
from . import kdtree  # noqa: F401
from .kdtree import (KDTree, ball_search, box_search, box_search_batch, build_kdtree,  # noqa: F401
                     deallocate_kdtree, knn_search)

__version__ = "0.1.0"
