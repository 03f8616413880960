"""Drop-in replacement of cosmokdtree's Python API on MI355X.

Same functions, arguments, defaults, assertions and return conventions as
src/cosmokdtree/kdtree.py of the reference:

    tree = build_kdtree(points, leaf=None, boxsize=None)
    deallocate_kdtree(tree)
    dist, idx = knn_search(tree, target, k, sorted=True)
    dist_list, idx_list = ball_search(tree, target, radius, sorted=False)
    idx = box_search(tree, box)

The work runs on the GPU through the C ABI of include/cosmokdtree_amd.h;
numpy inputs are copied to HBM and results copied back.  Differences from
the reference, all deliberate:

* Periodicity is chosen at run time: a ``boxsize`` with any entry other
  than -1 builds a periodic tree (the reference fixes it at compile time
  with PERIODIC and stops the process when a periodic build gets no box,
  cosmokdtree2py.f90:48-60).  Each tree keeps its own box size (the
  reference keeps one module-global L, cosmokdtree.f90:86-88,174).
* Points, targets, radii and boxes are converted to float64, as the
  reference does (src/cosmokdtree/kdtree.py:31,43,59,83), whatever their
  dtype, and distances come back as float64.  A float32 tree (what a
  DOUBLEPRECISION=0 Fortran build computes) is built only on request:
  ``build_kdtree(points, dtype=np.float32)``, or the device API
  (``_build`` with a float32 tensor); its distances are float32.
* Errors raise ValueError / MemoryError / RuntimeError instead of STOPping
  the process (cosmokdtree.f90:161,190; cosmokdtree2py.f90:58,276).
* Sorted results list equal distances in index order (the reference's
  quicksort leaves their order unspecified); unsorted kNN rows are returned
  in ascending order rather than heap order.

Device-resident work (no host copies) goes through the KDTree methods
``knn``, ``ball`` and ``box``, which take and return torch tensors on the
GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import TreeInfo, check, lib

__all__ = ["build_kdtree", "deallocate_kdtree", "knn_search", "ball_search", "box_search", "KDTree",
           "box_search_batch"]

_REGISTRY: dict[int, "KDTree"] = {}


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("cosmokdtree_amd needs a GPU (HIP device); none is visible")


class KDTree(int):
    """Handle of a tree in HBM.

    It is an ``int`` (the native tree address), as the reference's handle is
    (cosmokdtree2py.f90:70-73), and carries the tree's metadata.
    """

    def __new__(cls, ptr: int, **meta):
        obj = super().__new__(cls, ptr)
        obj.__dict__.update(meta)
        obj._ptr = C.c_void_p(ptr)
        obj._alive = True
        return obj

    # ----------------------------------------------------------- metadata
    def info(self) -> TreeInfo:
        self._check_alive()
        inf = TreeInfo()
        check(lib.ckd_get_info(self._ptr, C.byref(inf)), "ckd_get_info")
        return inf

    def _check_alive(self):
        if not self._alive:
            raise ValueError("the tree has been deallocated")

    def free(self):
        if self._alive:
            check(lib.ckd_free(self._ptr), "ckd_free")
            self._alive = False
            _REGISTRY.pop(int(self), None)

    def knn_config(self, kernel: str = "auto", stack: int = 0, queries_per_wave: int = 0):
        """Pick the kNN kernel ("auto", "lane", "wave", "stack"), the stack
        depth (0 = automatic, 32, 64) and the queries sharing a wave of the
        stack kernel (0 = 64, 1, 4, 16, 64).  Results do not depend on it."""
        self._check_alive()
        if kernel not in _lib.KNN_KERNELS:
            raise ValueError(f"kernel must be one of {sorted(_lib.KNN_KERNELS)}")
        check(lib.ckd_knn_config(self._ptr, _lib.KNN_KERNELS[kernel], int(stack), int(queries_per_wave)),
              "knn_config")
        return self

    def save(self, path: str):
        """Write the built tree to `path` (reload with kdtree.load_tree)."""
        self._check_alive()
        check(lib.ckd_tree_save(self._ptr, str(path).encode()), "save")

    def knn_timing(self, on: bool = True):
        """Start (on=True) or stop recording HIP events around this tree's kNN
        traversal kernel launches; returns (summed ms, launches) of the
        launches recorded since the previous call (synchronises them)."""
        self._check_alive()
        ms, n = C.c_double(0), C.c_int64(0)
        check(lib.ckd_knn_timing(self._ptr, int(bool(on)), C.byref(ms), C.byref(n)), "knn_timing")
        return ms.value, n.value

    # ------------------------------------------------------------ device API
    def _queries(self, q) -> torch.Tensor:
        q = torch.as_tensor(q, dtype=self.torch_dtype, device=self.device)
        if q.dim() == 1:
            q = q.reshape(1, -1)
        if q.dim() != 2 or q.shape[1] != self.ndim:
            raise ValueError(f"targets must have {self.ndim} columns, got shape {tuple(q.shape)}")
        return q.contiguous()

    def knn(self, q, k: int, sorted: bool = True, seed_r=None, stats=None):
        """kNN of device queries: returns (dist (nq,k), idx (nq,k) int64) on the GPU."""
        self._check_alive()
        q = self._queries(q)
        nq = q.shape[0]
        dist = torch.empty((nq, k), dtype=self.torch_dtype, device=self.device)
        idx = torch.empty((nq, k), dtype=torch.int64, device=self.device)
        if nq == 0 or k == 0:
            return dist, idx
        if seed_r is not None:
            seed_r = torch.as_tensor(seed_r, dtype=torch.float32, device=self.device).contiguous()
            if seed_r.numel() != nq:
                raise ValueError("seed_r needs one radius per query")
        check(lib.ckd_knn(self._ptr, _vp(q), nq, int(k), int(bool(sorted)), _vp(dist), _vp(idx), _vp(seed_r),
                          _vp(stats), _stream()), "knn_search")
        return dist, idx

    def sph_density(self, q, k: int = 32, mass=None):
        """SPH smoothing length and density of device queries, from their k
        nearest neighbours in one fused kernel (ckd_sph_density): returns
        (h (nq,), rho (nq,)) on the GPU with h = r_k / 2 and rho the cubic
        spline sum over the k neighbours.  mass: one per tree point in the
        original order, or None for unit masses.  3-D trees only."""
        self._check_alive()
        q = self._queries(q)
        nq = q.shape[0]
        h = torch.empty(nq, dtype=self.torch_dtype, device=self.device)
        rho = torch.empty(nq, dtype=self.torch_dtype, device=self.device)
        if mass is not None:
            mass = torch.as_tensor(mass, dtype=self.torch_dtype, device=self.device).contiguous()
            if mass.numel() != self.n:
                raise ValueError("mass needs one value per tree point")
        if nq:
            check(lib.ckd_sph_density(self._ptr, _vp(q), nq, int(k), _vp(mass), _vp(h), _vp(rho), _stream()),
                  "sph_density")
        return h, rho

    def ball(self, q, radius: float, sorted: bool = False):
        """Ball search of device queries: CSR (offsets (nq+1), dist, idx) on the GPU.

        One walk (ckd_ball_search) into buffers sized from the rows per query
        of this tree's previous ball search, scaled by (r / r_prev)^ndim;
        when the rows do not fit, a second walk (ckd_ball_fill) writes them
        into buffers of the exact size."""
        self._check_alive()
        q = self._queries(q)
        nq = q.shape[0]
        radius = float(radius)
        off = torch.empty(nq + 1, dtype=torch.int64, device=self.device)
        total = C.c_int64(0)
        cap = self._ball_capacity(nq, radius)
        dist = torch.empty(max(cap, 1), dtype=self.torch_dtype, device=self.device)
        idx = torch.empty(max(cap, 1), dtype=torch.int64, device=self.device)
        check(lib.ckd_ball_search(self._ptr, _vp(q), nq, radius, int(bool(sorted)), _vp(off), cap, _vp(dist),
                                  _vp(idx), C.byref(total), _stream()), "ball_search")
        tot = total.value
        if tot > cap:
            dist = torch.empty(max(tot, 1), dtype=self.torch_dtype, device=self.device)
            idx = torch.empty(max(tot, 1), dtype=torch.int64, device=self.device)
            check(lib.ckd_ball_fill(self._ptr, _vp(q), nq, radius, int(bool(sorted)), _vp(off), _vp(dist),
                                    _vp(idx), _stream()), "ball_search")
        if nq > 0 and radius > 0:
            self._ball_hint = (radius, tot / nq)
        return off, dist[:tot], idx[:tot]

    # rows per query assumed before a tree's first ball search, and the
    # largest first guess (rows); later calls use the previous call's count
    BALL_FIRST_ROWS_PER_QUERY = 64
    BALL_FIRST_MAX_ROWS = 1 << 26

    def _ball_capacity(self, nq: int, radius: float) -> int:
        hint = getattr(self, "_ball_hint", None)
        if hint is None:
            return int(min(nq * self.BALL_FIRST_ROWS_PER_QUERY, self.BALL_FIRST_MAX_ROWS))
        r0, per = hint
        per = per * (radius / r0) ** self.ndim if r0 > 0 else per
        return int(min(nq * per * 1.25 + 4096, 1 << 40))

    def box(self, boxes, wrap: bool = False):
        """Box search of device boxes (nb, 2*ndim): CSR (offsets, idx) on the GPU.

        wrap=True (periodic trees) takes each box modulo the periodic box, so
        boxes may straddle the faces (ckd_box_count_periodic); the default
        follows the reference, which does not wrap (cosmokdtree.f90:1260)."""
        self._check_alive()
        b = torch.as_tensor(boxes, dtype=self.torch_dtype, device=self.device)
        if b.dim() == 1:
            b = b.reshape(1, -1)
        if b.dim() != 2 or b.shape[1] != 2 * self.ndim:
            raise ValueError("Error: box size does not match tree dimensionality.")
        b = b.contiguous()
        nb = b.shape[0]
        off = torch.empty(nb + 1, dtype=torch.int64, device=self.device)
        total = C.c_int64(0)
        count, fill = ((lib.ckd_box_count_periodic, lib.ckd_box_fill_periodic) if wrap
                       else (lib.ckd_box_count, lib.ckd_box_fill))
        check(count(self._ptr, _vp(b), nb, _vp(off), C.byref(total), _stream()), "box_search")
        idx = torch.empty(max(total.value, 1), dtype=torch.int64, device=self.device)
        check(fill(self._ptr, _vp(b), nb, _vp(off), _vp(idx), _stream()), "box_search")
        return off, idx[:total.value]

    def dump(self) -> dict:
        """Structure arrays on the host (library node numbering)."""
        inf = self.info()
        nn, n, d = inf.nnodes, inf.n, inf.ndim
        dt = self.np_dtype
        out = dict(axis=np.empty(nn, np.int32), split=np.empty(nn, dt), lo=np.empty((nn, d), dt),
                   hi=np.empty((nn, d), dt), child=np.empty(nn, np.int32), parent=np.empty(nn, np.int32),
                   start=np.empty(nn, np.int64), count=np.empty(nn, np.int64), points=np.empty((n, d), dt),
                   idx=np.empty(n, np.int64))
        torch.cuda.synchronize()
        check(lib.ckd_tree_dump(self._ptr, *[out[k].ctypes.data_as(C.c_void_p) for k in (
            "axis", "split", "lo", "hi", "child", "parent", "start", "count", "points", "idx")]), "dump")
        return out


def _tree(tree) -> KDTree:
    if isinstance(tree, KDTree):
        tree._check_alive()
        return tree
    t = _REGISTRY.get(int(tree))
    if t is None:
        raise ValueError("not a tree handle returned by build_kdtree")
    return t


def _build(points, leaf: int, boxsize, keep_points: bool = False, idx_bits: int = 0) -> KDTree:
    """Build from a torch tensor or array-like; boxsize None = non-periodic.
    keep_points keeps the device copy of the input on the handle (.points);
    idx_bits 32 or 64 fixes the stored index width (0: 32 below 2^31
    points, the reference's LONGINT=1 is 64)."""
    _require_gpu()
    if isinstance(points, torch.Tensor):
        dt = torch.float32 if points.dtype == torch.float32 else torch.float64
        pts = points.to(device="cuda", dtype=dt)
    else:
        arr = np.asarray(points)
        dt = torch.float32 if arr.dtype == np.float32 else torch.float64
        pts = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float32 if dt == torch.float32 else np.float64))
        pts = pts.to("cuda")
    if pts.dim() != 2:
        raise ValueError("points must be a 2-D array (npoints, ndim)")
    pts = pts.contiguous()
    n, d = pts.shape
    bs = None
    if boxsize is not None:
        bs = (C.c_double * d)(*[float(x) for x in boxsize])
    out = C.c_void_p()
    prec = 4 if dt == torch.float32 else 8
    check(lib.ckd_build(_vp(pts), n, d, prec, int(leaf), bs, int(idx_bits), _stream(), C.byref(out)), "build_kdtree")
    tree = KDTree(out.value, n=n, ndim=d, leafsize=int(leaf), torch_dtype=dt,
                  np_dtype=np.float32 if prec == 4 else np.float64, device=pts.device,
                  periodic=boxsize is not None,
                  boxsize=None if boxsize is None else np.asarray(boxsize, dtype=np.float64),
                  points=pts if keep_points else None)
    _REGISTRY[int(tree)] = tree
    return tree


def load_tree(path: str) -> KDTree:
    """A tree saved by KDTree.save, on the current device, ready to query."""
    _require_gpu()
    out = C.c_void_p()
    check(lib.ckd_tree_load(str(path).encode(), _stream(), C.byref(out)), "load")
    inf = TreeInfo()
    check(lib.ckd_get_info(out, C.byref(inf)), "ckd_get_info")
    prec, d = inf.prec, inf.ndim
    dt = torch.float32 if prec == 4 else torch.float64
    bs = np.array(inf.boxsize[:d], dtype=np.float64) if inf.periodic else None
    tree = KDTree(out.value, n=inf.n, ndim=d, leafsize=inf.leafsize, torch_dtype=dt,
                  np_dtype=np.float32 if prec == 4 else np.float64, device=torch.device("cuda", torch.cuda.current_device()),
                  periodic=bool(inf.periodic), boxsize=bs, points=None)
    _REGISTRY[int(tree)] = tree
    return tree


def release_memory() -> None:
    """Return the library's cached but unused device memory to the driver.

    Builds and queries take their device memory from a per-device cache
    that keeps freed blocks for the next call; live trees keep theirs.  Call
    this before a phase that needs the HBM elsewhere (torch's caching
    allocator does not see the library's cache).  Waits for all work on the
    devices used."""
    check(lib.ckd_release_memory(), "release memory")


# --------------------------------------------------------------- drop-in API

def build_kdtree(points, leaf=None, boxsize=None, *, dtype=np.float64):
    """src/cosmokdtree/kdtree.py:16-34.  The points are converted to float64
    (kdtree.py:31) unless the keyword dtype=np.float32 asks for a float32
    tree (no reference counterpart: the reference Python path is float64
    only, src/Makefile:22-25)."""
    dtype = np.dtype(dtype)
    if dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
        raise ValueError("dtype must be float64 (the reference's) or float32")
    if leaf is None:
        leaf = 16
    assert leaf > 0, "Leaf size must be a positive integer"
    if boxsize is None:
        boxsize = -1. * np.ones(points.shape[1])
    assert len(boxsize) == points.shape[1], "Incorrect dimensions for boxsize var. Size must be [Lx, Ly, Lz, ...]"
    boxsize = np.array(boxsize, dtype=np.float64)
    periodic = bool(np.any(boxsize != -1.))  # cosmokdtree2py.f90:48-55
    if isinstance(points, torch.Tensor):
        points = points.detach().cpu().numpy()
    points = np.ascontiguousarray(points, dtype=dtype)
    return _build(points, int(leaf), boxsize if periodic else None)


def deallocate_kdtree(tree):
    """src/cosmokdtree/kdtree.py:37-38."""
    _tree(tree).free()


def _host_targets(t: KDTree, target) -> np.ndarray:
    target = np.array(target, dtype=t.np_dtype)
    if len(target.shape) == 1:
        target = target.reshape(1, -1)
    if target.shape[1] != t.ndim:
        raise ValueError(f"targets must have {t.ndim} columns, got {target.shape[1]}")
    return target


def knn_search(tree, target, k, sorted=True):
    """src/cosmokdtree/kdtree.py:41-53: (dist (ntar,k), idx (ntar,k)), 0-based,
    empty slots -2."""
    assert type(k) in [int, np.int32, np.int64], "k must be an integer"
    t = _tree(tree)
    target = _host_targets(t, target)
    if k < 0:
        raise ValueError("k must not be negative")
    dist, idx = t.knn(torch.from_numpy(target), int(k), sorted=bool(sorted))
    dist = dist.cpu().numpy()
    idx = idx.cpu().numpy()
    return np.asfortranarray(dist), np.asfortranarray(idx)


def ball_search(tree, target, radius, sorted=False):
    """src/cosmokdtree/kdtree.py:56-79: per-target lists of distances and indices."""
    assert radius > 0., "Radius must be a positive number"
    radius = float(radius)
    t = _tree(tree)
    target = _host_targets(t, target)
    off, dist, idx = t.ball(torch.from_numpy(target), radius, sorted=bool(sorted))
    off = off.cpu().numpy()
    dist = dist.cpu().numpy()
    idx = idx.cpu().numpy()
    ntar = target.shape[0]
    return ([dist[off[i]:off[i + 1]] for i in range(ntar)], [idx[off[i]:off[i + 1]] for i in range(ntar)])


def box_search(tree, box):
    """src/cosmokdtree/kdtree.py:82-87: indices (0-based) of the points in one
    box (xmin, xmax, ymin, ymax, ...)."""
    box = np.array(box, dtype=np.float64)
    t = _tree(tree)
    if box.size != 2 * t.ndim:  # cosmokdtree2py.f90:275-277
        raise ValueError("Error: box size does not match tree dimensionality.")
    off, idx = t.box(torch.from_numpy(box.reshape(1, -1).astype(t.np_dtype)))
    return idx.cpu().numpy()


def box_search_batch(tree, boxes, wrap=False):
    """Many boxes in one call (no reference counterpart: the reference takes
    one box per call).  Returns CSR (offsets (nb+1), idx) as numpy arrays.
    wrap=True takes the boxes modulo a periodic tree's box (KDTree.box)."""
    t = _tree(tree)
    boxes = np.asarray(np.asarray(boxes, dtype=np.float64), dtype=t.np_dtype).reshape(-1, 2 * t.ndim)
    off, idx = t.box(torch.from_numpy(boxes), wrap=bool(wrap))
    return off.cpu().numpy(), idx.cpu().numpy()
