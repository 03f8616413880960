"""CPU checkers for the cosmokdtree build and query path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module; the product
package ``cosmokdtree_amd`` never does.

Two checkers live here:

``OracleTree``
    ctypes wrapper of ``libckd_oracle.so`` (``ckd_oracle.c``), a C
    restatement of the reference routines in src/cosmokdtree.f90 and
    src/cosmokdtree2py.f90.  Its outputs follow src/cosmokdtree/kdtree.py:
    0-based indices (an empty kNN slot reads -2), distances after sqrt,
    ball results as per-target lists.

``RefTree``
    ctypes wrapper of ``_ref/libckdref_<variant>.so``: the reference's own
    src/cosmokdtree.f90 compiled by ``oracle/Makefile`` with a thin
    bind(C) harness (``ref_bridge.f90``) that mirrors the f2py bridge.
    Built only where the reference sources are present; tests skip it
    elsewhere.

Both expose the same methods (``knn``, ``ball``, ``box``, ``dump``) so a test
can run either against the same inputs.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CKD_ORACLE_SO selects another build of the same source (the sanitizer build)
ORACLE_SO = os.environ.get("CKD_ORACLE_SO") or os.path.join(HERE, "libckd_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_vp)


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built (run make -C oracle)")
        lib = C.CDLL(ORACLE_SO)
        lib.ckdo_build.restype = _vp
        lib.ckdo_build.argtypes = [_vp, _i64, _i32, _i32, _i32, _vp, C.POINTER(_i32)]
        lib.ckdo_free.argtypes = [_vp]
        lib.ckdo_nnodes.restype = _i64
        lib.ckdo_nnodes.argtypes = [_vp]
        lib.ckdo_dump.argtypes = [_vp] + [_vp] * 11
        lib.ckdo_knn.argtypes = [_vp, _vp, _i64, _i32, _i32, _vp, _vp]
        lib.ckdo_ball_run.restype = _vp
        lib.ckdo_ball_run.argtypes = [_vp, _vp, _i64, C.c_double, _i32, _vp]
        lib.ckdo_box_run.restype = _vp
        lib.ckdo_box_run.argtypes = [_vp, _vp, _i64, _vp]
        lib.ckdo_box_wrap_run.restype = _vp
        lib.ckdo_box_wrap_run.argtypes = [_vp, _vp, _i64, _vp]
        lib.ckdo_res_fetch.argtypes = [_vp, _vp, _vp, _vp]
        lib.ckdo_res_free.argtypes = [_vp]
        _oracle = lib
    return _oracle


_ERRORS = {
    1: "points must have between 1 and 8 columns",
    2: "Leaf size must be a positive integer",
    3: "Points outside (-L/2, L/2) range !!",
    4: "at least one point is required",
    5: "boxsize entries must be positive",
}


def _csr(counts):
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    return off


def _split(off, arr):
    return [arr[off[i]:off[i + 1]] for i in range(len(off) - 1)]


class OracleTree:
    """C restatement of the reference; see module docstring."""

    def __init__(self, points, leaf=16, boxsize=None):
        pts = np.asarray(points)
        if pts.dtype not in (np.float32, np.float64):
            pts = pts.astype(np.float64)
        self.dtype = pts.dtype
        self.prec = 4 if pts.dtype == np.float32 else 8
        pts = np.ascontiguousarray(pts)
        self.n, self.ndim = pts.shape
        L = None if boxsize is None else np.ascontiguousarray(np.asarray(boxsize, dtype=self.dtype))
        self.periodic = L is not None
        self.L = L
        err = _i32(0)
        self.lib = oracle_lib()
        self.h = self.lib.ckdo_build(_ptr(pts), self.n, self.ndim, self.prec, int(leaf), _ptr(L), C.byref(err))
        if not self.h:
            raise ValueError(_ERRORS.get(err.value, f"oracle build error {err.value}"))

    def close(self):
        if self.h:
            self.lib.ckdo_free(self.h)
            self.h = None

    __del__ = close

    def _q(self, q):
        q = np.asarray(q, dtype=self.dtype)
        if q.ndim == 1:
            q = q.reshape(1, -1)
        return np.ascontiguousarray(q)

    def knn(self, q, k, sorted=True):
        q = self._q(q)
        nq = q.shape[0]
        d = np.empty((nq, k), dtype=self.dtype)
        i = np.empty((nq, k), dtype=np.int64)
        self.lib.ckdo_knn(self.h, _ptr(q), nq, int(k), int(bool(sorted)), _ptr(d), _ptr(i))
        return d, i

    def ball_csr(self, q, r, sorted=False):
        q = self._q(q)
        nq = q.shape[0]
        cnt = np.empty(nq, dtype=np.int64)
        res = self.lib.ckdo_ball_run(self.h, _ptr(q), nq, float(r), int(bool(sorted)), _ptr(cnt))
        off = _csr(cnt)
        d = np.empty(max(int(off[-1]), 1), dtype=self.dtype)
        i = np.empty(max(int(off[-1]), 1), dtype=np.int64)
        self.lib.ckdo_res_fetch(res, _ptr(off), _ptr(d), _ptr(i))
        self.lib.ckdo_res_free(res)
        return off, d[:off[-1]], i[:off[-1]]

    def ball(self, q, r, sorted=False):
        off, d, i = self.ball_csr(q, r, sorted)
        return _split(off, d), _split(off, i)

    def box_csr(self, boxes, wrap=False):
        """Box search; wrap=True takes boxes modulo the periodic box (the
        straddling-box rule of ckdo_box_wrap_run)."""
        b = np.ascontiguousarray(np.asarray(boxes, dtype=self.dtype).reshape(-1, 2 * self.ndim))
        nb = b.shape[0]
        cnt = np.empty(nb, dtype=np.int64)
        if wrap and not self.periodic:
            raise ValueError("wrapped boxes need a periodic tree")
        run = self.lib.ckdo_box_wrap_run if wrap else self.lib.ckdo_box_run
        res = run(self.h, _ptr(b), nb, _ptr(cnt))
        off = _csr(cnt)
        i = np.empty(max(int(off[-1]), 1), dtype=np.int64)
        self.lib.ckdo_res_fetch(res, _ptr(off), None, _ptr(i))
        self.lib.ckdo_res_free(res)
        return off, i[:off[-1]]

    def box(self, box):
        off, i = self.box_csr(box)
        return i

    def dump(self):
        """Preorder structure: dict of node arrays + leaf index pool."""
        nn = int(self.lib.ckdo_nnodes(self.h))
        out = dict(
            is_leaf=np.empty(nn, np.int8), axis=np.empty(nn, np.int32), split=np.empty(nn, self.dtype),
            minb=np.empty((nn, self.ndim), self.dtype), maxb=np.empty((nn, self.ndim), self.dtype),
            left=np.empty(nn, np.int64), right=np.empty(nn, np.int64), start=np.empty(nn, np.int64),
            count=np.empty(nn, np.int64), lpts=np.empty((self.n, self.ndim), self.dtype),
            lidx=np.empty(self.n, np.int64))
        self.lib.ckdo_dump(self.h, *[_ptr(out[k]) for k in (
            "is_leaf", "axis", "split", "minb", "maxb", "left", "right", "start", "count", "lpts", "lidx")])
        return out


# ---------------------------------------------------------------- reference

def ref_lib_path(periodic, ndim, dtype):
    bits = 32 if np.dtype(dtype) == np.float32 else 64
    return os.path.join(REF_DIR, f"libckdref_p{int(bool(periodic))}_d{ndim}_f{bits}.so")


def ref_available(periodic=False, ndim=3, dtype=np.float64):
    return os.path.exists(ref_lib_path(periodic, ndim, dtype))


_ref_libs = {}


def _ref_lib(path):
    lib = _ref_libs.get(path)
    if lib is None:
        lib = C.CDLL(path)
        lib.ckdr_build.restype = _vp
        lib.ckdr_build.argtypes = [_vp, _i64, _i32, _vp]
        lib.ckdr_free.argtypes = [_vp]
        lib.ckdr_knn.argtypes = [_vp, _vp, _i64, _i32, _i32, _vp, _vp]
        lib.ckdr_ball_run.argtypes = [_vp, _vp, _i64, C.c_double, _i32, _vp]
        lib.ckdr_box_run.argtypes = [_vp, _vp, _i64, _vp]
        lib.ckdr_res_fetch.argtypes = [_i64, _vp, _vp, _vp]
        lib.ckdr_nnodes.restype = _i64
        lib.ckdr_nnodes.argtypes = [_vp]
        lib.ckdr_dump.argtypes = [_vp, _i64, _i64] + [_vp] * 10
        _ref_libs[path] = lib
    return lib


class RefTree:
    """The reference Fortran module (compiled here) behind the same methods.

    Mirrors src/cosmokdtree/kdtree.py: indices are returned 0-based (the
    harness returns them 1-based, as the f2py bridge does, and this wrapper
    subtracts 1 as kdtree.py:52 does).  The Fortran module STOPs the process
    on invalid input (cosmokdtree.f90:161,190), so callers validate first.
    """

    def __init__(self, points, leaf=16, boxsize=None):
        pts = np.asarray(points)
        self.dtype = np.float32 if pts.dtype == np.float32 else np.float64
        pts = np.asfortranarray(pts, dtype=self.dtype)
        self.n, self.ndim = pts.shape
        self.periodic = boxsize is not None
        path = ref_lib_path(self.periodic, self.ndim, self.dtype)
        if not os.path.exists(path):
            raise RuntimeError(f"reference library {path} is not built")
        self.lib = _ref_lib(path)
        L = None
        if self.periodic:
            L = np.ascontiguousarray(np.asarray(boxsize, dtype=self.dtype))
            if np.any(pts < -L / 2) or np.any(pts > L / 2):
                raise ValueError("Points outside (-L/2, L/2) range !!")
        self._keep = (pts, L)
        self.h = self.lib.ckdr_build(_ptr(pts), self.n, int(leaf), _ptr(L))

    def close(self):
        if self.h:
            self.lib.ckdr_free(self.h)
            self.h = None

    __del__ = close

    def _q(self, q):
        q = np.asarray(q, dtype=self.dtype)
        if q.ndim == 1:
            q = q.reshape(1, -1)
        return np.asfortranarray(q)

    def knn(self, q, k, sorted=True):
        q = self._q(q)
        nq = q.shape[0]
        d = np.empty((nq, k), dtype=self.dtype, order="F")
        i = np.empty((nq, k), dtype=np.int64, order="F")
        self.lib.ckdr_knn(self.h, _ptr(q), nq, int(k), int(bool(sorted)), _ptr(d), _ptr(i))
        i -= 1
        return np.ascontiguousarray(d), np.ascontiguousarray(i)

    def _fetch(self, cnt, with_dist):
        off = _csr(cnt)
        tot = max(int(off[-1]), 1)
        i = np.empty(tot, dtype=np.int64)
        d = np.empty(tot, dtype=self.dtype) if with_dist else None
        self.lib.ckdr_res_fetch(len(cnt), _ptr(off), _ptr(d), _ptr(i))
        i -= 1
        i = i[:off[-1]]
        return off, (d[:off[-1]] if with_dist else None), i

    def ball_csr(self, q, r, sorted=False):
        q = self._q(q)
        nq = q.shape[0]
        cnt = np.empty(nq, dtype=np.int64)
        self.lib.ckdr_ball_run(self.h, _ptr(q), nq, float(r), int(bool(sorted)), _ptr(cnt))
        return self._fetch(cnt, True)

    def ball(self, q, r, sorted=False):
        off, d, i = self.ball_csr(q, r, sorted)
        return _split(off, d), _split(off, i)

    def box_csr(self, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, dtype=self.dtype).reshape(-1, 2 * self.ndim))
        nb = b.shape[0]
        cnt = np.empty(nb, dtype=np.int64)
        self.lib.ckdr_box_run(self.h, _ptr(b), nb, _ptr(cnt))
        off, _, i = self._fetch(cnt, False)
        return off, i

    def box(self, box):
        return self.box_csr(box)[1]

    def dump(self):
        nn = int(self.lib.ckdr_nnodes(self.h))
        out = dict(
            is_leaf=np.empty(nn, np.int8), axis=np.empty(nn, np.int32), split=np.empty(nn, self.dtype),
            minb=np.empty((nn, self.ndim), self.dtype), maxb=np.empty((nn, self.ndim), self.dtype),
            left=np.empty(nn, np.int64), right=np.empty(nn, np.int64), start=np.empty(nn, np.int64),
            count=np.empty(nn, np.int64), lidx=np.empty(self.n, np.int64))
        self.lib.ckdr_dump(self.h, nn, self.n, *[_ptr(out[k]) for k in (
            "is_leaf", "axis", "split", "minb", "maxb", "left", "right", "start", "count", "lidx")])
        return out
