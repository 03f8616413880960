"""The f2py-level interface of the reference, on the GPU library.

The reference's Python wrapper (src/cosmokdtree/kdtree.py) calls an f2py
extension module, ``pycosmokdtree``, whose functions live in the Fortran
module of the same name: ``pycosmokdtree.pycosmokdtree.<function>``
(src/cosmokdtree2py.f90:18-19,29-315).  This module exposes the same seven
functions with the same arguments and return conventions, so that wrapper
(or any other caller of the f2py module) runs unchanged on the GPU:

    import cosmokdtree_amd.pycosmokdtree as pycosmokdtree
    tree = pycosmokdtree.pycosmokdtree.pybuild_kdtree(points, 16, boxsize)

Conventions kept from the bridge:

* points and targets are float64; the tree is a float64 tree with int64
  handle (the C address of the tree, an ``int``);
* indices are 1-based, an empty kNN slot is -1 (the wrapper's ``idx -= 1``
  then gives 0-based indices and -2, src/cosmokdtree/kdtree.py:52);
* outputs are Fortran-ordered numpy arrays; the ball counts are int32;
* ball and box search are two calls: the first returns counts and an opaque
  per-target handle array, the second the (zero-padded) rows, and releases
  the pending result (cosmokdtree2py.f90:153-315).

Periodicity is chosen from ``boxsize`` at run time: any entry other than -1
makes a periodic tree (the reference chooses it at compile time,
cosmokdtree2py.f90:48-67).
"""
from __future__ import annotations

import itertools

import numpy as np
import torch

from .kdtree import _build, _tree

_pending: dict[int, tuple] = {}  # results between the two ball / box calls
_keys = itertools.count(1)


def _targets(targett, ndim) -> torch.Tensor:
    t = np.ascontiguousarray(np.asarray(targett, dtype=np.float64)).reshape(-1, int(ndim))
    return torch.from_numpy(t)


class _Module:
    """The functions of the Fortran module ``pycosmokdtree``."""

    @staticmethod
    def pybuild_kdtree(points, leaf, boxsize):
        """cosmokdtree2py.f90:29-74: returns the tree handle."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
        bs = np.asarray(boxsize, dtype=np.float64)
        periodic = bool(np.any(bs != -1.0))
        return int(_build(pts, int(leaf), bs if periodic else None))

    @staticmethod
    def pydeallocate_tree(tree):
        """cosmokdtree2py.f90:80-97."""
        _tree(tree).free()

    @staticmethod
    def pyknn_search(tree, targett, ndim, ntar, k, sorted):
        """cosmokdtree2py.f90:103-143: (dist (ntar, k), idx (ntar, k) 1-based)."""
        t = _tree(tree)
        d, i = t.knn(_targets(targett, ndim), int(k), bool(sorted))
        i = i.cpu().numpy()
        i = np.where(i >= 0, i + 1, -1)
        return np.asfortranarray(d.cpu().numpy()), np.asfortranarray(i)

    @staticmethod
    def pyball_search_call_1(tree, targett, ndim, ntar, radius, sorted):
        """cosmokdtree2py.f90:153-201: (nball int32 (ntar), nballmax, query)."""
        off, d, i = _tree(tree).ball(_targets(targett, ndim), float(radius), bool(sorted))
        off = off.cpu().numpy()
        nball = np.diff(off).astype(np.int32)
        key = next(_keys)
        _pending[key] = (off, d.cpu().numpy(), i.cpu().numpy() + 1)
        return nball, int(nball.max(initial=0)), np.full(int(ntar), key, dtype=np.int64)

    @staticmethod
    def pyball_search_call_2(ndim, ntar, c_ptr_query, nballmax, nball):
        """cosmokdtree2py.f90:203-246: zero-padded (ntar, nballmax) rows."""
        key = int(np.asarray(c_ptr_query).ravel()[0]) if int(ntar) > 0 else None
        off, d, i = _pending.pop(key) if key is not None else (np.zeros(1, np.int64), None, None)
        nb = np.asarray(nball)
        dist = np.zeros((int(ntar), int(nballmax)), dtype=np.float64, order="F")
        idx = np.zeros((int(ntar), int(nballmax)), dtype=np.int64, order="F")
        for r in range(int(ntar)):
            dist[r, :nb[r]] = d[off[r]:off[r] + nb[r]]
            idx[r, :nb[r]] = i[off[r]:off[r] + nb[r]]
        return dist, idx

    @staticmethod
    def pybox_search_call_1(tree, box):
        """cosmokdtree2py.f90:252-287: (nbox, query)."""
        t = _tree(tree)
        b = np.asarray(box, dtype=np.float64).ravel()
        if b.size != 2 * t.ndim:
            raise ValueError("box must hold 2 * ndim values (xmin, xmax, ymin, ymax, ...)")
        off, i = t.box(torch.from_numpy(b.reshape(1, -1)))
        key = next(_keys)
        _pending[key] = (i.cpu().numpy() + 1,)
        return int(off[-1].item()), key

    @staticmethod
    def pybox_search_call_2(c_ptr_query, nbox):
        """cosmokdtree2py.f90:289-315: idx (nbox) 1-based."""
        (idx,) = _pending.pop(int(c_ptr_query))
        return idx[:int(nbox)]


pycosmokdtree = _Module()
