"""Batched queries over many independent point sets (a forest).

    forest = build_forest([pts_0, pts_1, ...], leaf=16, boxsize=[10] * 3)
    plan = BatchPlan(forest, queries_per_set, k_per_set, radius_per_set)
    out = plan.run(queries)          # eager
    plan.capture(queries)            # record the same launches in a hipGraph
    out = plan.replay()              # replay the graph

All sets are built in one level-synchronous pass (ckd_build_forest) and
all queries of all sets, with their own k and ball radius, are answered by
persistent kernels that drain one device work queue (ckd_batch_*).  Each
set's results equal those of the single-tree calls on that set; indices are
relative to the set.  No reference counterpart: the reference answers one
tree per call (src/cosmokdtree/kdtree.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib
from .kdtree import KDTree, _REGISTRY, _require_gpu, _stream, _vp

_vp_t = C.c_void_p
lib.ckd_batch_plan.argtypes = [_vp_t, _vp_t, _vp_t, _vp_t, C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                               C.POINTER(_vp_t), C.POINTER(C.c_int64)]
lib.ckd_batch_run.argtypes = [_vp_t] * 8
lib.ckd_batch_status.argtypes = [_vp_t, _vp_t, C.POINTER(C.c_int64), C.POINTER(C.c_int32), _vp_t]
lib.ckd_batch_free.argtypes = [_vp_t]
lib.ckd_batch_config.argtypes = [_vp_t, C.c_int32, C.c_int32, C.c_int32]
lib.ckd_batch_config_ball.argtypes = [_vp_t, C.c_int32]
for _n in ("ckd_batch_plan", "ckd_batch_run", "ckd_batch_status", "ckd_batch_free", "ckd_batch_config",
           "ckd_batch_config_ball"):
    getattr(lib, _n).restype = C.c_int


def build_forest(sets, leaf: int = 16, boxsize=None, dtype=None, idx_bits: int = 0) -> KDTree:
    """One tree per point set, built in one pass.  `sets` is a list of (n_i, 3)
    arrays/tensors, or a pair (points (N, 3), offsets (nsets+1)).  idx_bits
    32 or 64 fixes the stored index width (0: 32 below 2^31 points)."""
    _require_gpu()
    if isinstance(sets, tuple) and len(sets) == 2:
        pts, off = sets
        off = np.asarray(off, dtype=np.int64)
        pts = torch.as_tensor(pts)
    else:
        arrs = [torch.as_tensor(np.asarray(s) if not isinstance(s, torch.Tensor) else s) for s in sets]
        off = np.zeros(len(arrs) + 1, dtype=np.int64)
        np.cumsum([a.shape[0] for a in arrs], out=off[1:])
        pts = torch.cat([a.to("cuda") for a in arrs]) if arrs else torch.empty((0, 3))
    dt = dtype or (torch.float32 if pts.dtype == torch.float32 else torch.float64)
    pts = pts.to(device="cuda", dtype=dt).contiguous()
    n, d = pts.shape
    bs = None if boxsize is None else (C.c_double * d)(*[float(x) for x in boxsize])
    out = C.c_void_p()
    prec = 4 if dt == torch.float32 else 8
    check(lib.ckd_build_forest(_vp(pts), off.ctypes.data_as(C.c_void_p), len(off) - 1, d, prec, int(leaf), bs,
                               int(idx_bits), _stream(), C.byref(out)), "build_forest")
    tree = KDTree(out.value, n=n, ndim=d, leafsize=int(leaf), torch_dtype=dt,
                  np_dtype=np.float32 if prec == 4 else np.float64, device=pts.device,
                  periodic=boxsize is not None,
                  boxsize=None if boxsize is None else np.asarray(boxsize, dtype=np.float64), points=None,
                  set_offsets=off)
    _REGISTRY[int(tree)] = tree
    return tree


INT32_MAX = 2 ** 31 - 1


def check_plan_limits(n_points: int, qoff) -> None:
    """The batch kernels hold rows, query positions and leaf ranks in 32
    bits: forests and query batches must stay below 2^31 - 1 rows (the C
    plan rejects them too, with CKD_ERR_LIMIT)."""
    qoff = np.asarray(qoff, dtype=np.int64)
    if n_points >= INT32_MAX:
        raise ValueError(f"the batched engine takes forests below 2^31 - 1 points (got {n_points})")
    if len(qoff) and qoff[-1] >= INT32_MAX:
        raise ValueError(f"the batched engine takes batches below 2^31 - 1 queries (got {int(qoff[-1])})")
    if len(qoff) and (qoff[0] != 0 or np.any(np.diff(qoff) < 0)):
        raise ValueError("query counts must not be negative")


class BatchPlan:
    """Work queue and output layout for one batched query workload."""

    def __init__(self, forest: KDTree, nq_per_set, k_per_set, radius_per_set=None, sorted: bool = True,
                 threads: int = 256, blocks: int = 0, ball_capacity: int | None = None, knn_kernel: str | None = None,
                 queries_per_wave: int = 64, stack_depth: int = 0, ball_kernel: str | None = None):
        self.forest = forest
        self.ball_kernel = ball_kernel
        self.knn_kernel, self.queries_per_wave, self.stack_depth = knn_kernel, int(queries_per_wave), int(stack_depth)
        nq = np.asarray(nq_per_set, dtype=np.int64)
        self.qoff = np.zeros(len(nq) + 1, dtype=np.int64)
        np.cumsum(nq, out=self.qoff[1:])
        self.k = np.ascontiguousarray(np.asarray(k_per_set, dtype=np.int32))
        self.radius = None if radius_per_set is None else np.ascontiguousarray(
            np.asarray(radius_per_set, dtype=np.float64))
        self.nq = int(self.qoff[-1])
        check_plan_limits(forest.n, self.qoff)
        # a set takes part in the ball search only with a positive radius
        self.has_ball = self.radius is not None and bool(np.any(self.radius > 0))
        self.sorted = bool(sorted)
        self.threads, self.blocks = int(threads), int(blocks)
        self.ball_capacity = ball_capacity
        self._plan = None
        self._graph = None
        self._make(ball_capacity or 0)

    def _make(self, capacity):
        # a captured graph holds the old plan's buffers: it must not run again
        self._graph = None
        if self._plan:
            lib.ckd_batch_free(self._plan)
            self._plan = None
        plan = C.c_void_p()
        tot = C.c_int64(0)
        check(lib.ckd_batch_plan(self.forest._ptr, self.qoff.ctypes.data_as(C.c_void_p),
                                 self.k.ctypes.data_as(C.c_void_p),
                                 None if self.radius is None else self.radius.ctypes.data_as(C.c_void_p),
                                 int(self.sorted), self.threads, self.blocks, int(capacity), C.byref(plan),
                                 C.byref(tot)), "batch_plan")
        self._plan = plan
        self._apply_config()
        self.knn_total = tot.value
        self.capacity = int(capacity)
        dev, dt = self.forest.device, self.forest.torch_dtype
        self.knn_dist = torch.empty(max(self.knn_total, 1), dtype=dt, device=dev)
        self.knn_idx = torch.empty(max(self.knn_total, 1), dtype=torch.int64, device=dev)
        # zeros: the kernels write the offsets only when some set has a ball search
        self.ball_off = torch.zeros(self.nq + 1, dtype=torch.int64, device=dev)
        self.ball_dist = torch.empty(max(self.capacity, 1), dtype=dt, device=dev)
        self.ball_idx = torch.empty(max(self.capacity, 1), dtype=torch.int64, device=dev)

    def _apply_config(self):
        kinds = {"lane": 0, "stack": 2}
        if self.ball_kernel is not None:  # None: the plan's own default (by query density)
            if self.ball_kernel not in kinds:
                raise ValueError("ball_kernel must be 'lane' or 'stack'")
            check(lib.ckd_batch_config_ball(self._plan, kinds[self.ball_kernel]), "batch_config_ball")
        if self.knn_kernel is None:  # the plan's own default (by query density)
            return
        if self.knn_kernel not in kinds:
            raise ValueError("knn_kernel must be 'lane' or 'stack'")
        check(lib.ckd_batch_config(self._plan, kinds[self.knn_kernel], self.queries_per_wave, self.stack_depth),
              "batch_config")

    def config(self, knn_kernel: str = "stack", queries_per_wave: int = 64, stack_depth: int = 0,
               ball_kernel: str | None = None):
        """kNN kernel ("stack": shared walks, queries in leaf order, or
        "lane": one walk per query), queries per wave of the stack walk
        (1, 4, 16, 64) and its stack depth (0, 32, 64); ball kernel ("lane"
        or "stack", the shared walk).  Results do not depend on it."""
        self.knn_kernel, self.queries_per_wave, self.stack_depth = knn_kernel, int(queries_per_wave), int(stack_depth)
        if ball_kernel is not None:
            self.ball_kernel = ball_kernel
        self._apply_config()
        self._graph = None
        return self

    def __del__(self):
        if getattr(self, "_plan", None):
            lib.ckd_batch_free(self._plan)
            self._plan = None

    def launch(self, queries: torch.Tensor):
        """Enqueue the batched kernels on the current stream (capturable)."""
        self._q = queries
        check(lib.ckd_batch_run(self._plan, _vp(queries), _vp(self.knn_dist), _vp(self.knn_idx),
                                _vp(self.ball_off), _vp(self.ball_dist), _vp(self.ball_idx), _stream()),
              "batch_run")

    def status(self):
        tot, ovf = C.c_int64(0), C.c_int32(0)
        check(lib.ckd_batch_status(self._plan, _vp(self.ball_off), C.byref(tot), C.byref(ovf), _stream()),
              "batch_status")
        return tot.value, bool(ovf.value)

    def run(self, queries: torch.Tensor):
        """Eager run; grows the ball capacity and reruns once if it was short."""
        queries = queries.to(device=self.forest.device, dtype=self.forest.torch_dtype).contiguous()
        self.launch(queries)
        if self.has_ball:
            tot, ovf = self.status()
            if ovf:
                self._make(tot)
                self.launch(queries)
        return self.results()

    def results(self):
        tot = int(self.ball_off[-1].item()) if self.has_ball else 0
        return dict(knn_dist=self.knn_dist[:self.knn_total], knn_idx=self.knn_idx[:self.knn_total],
                    ball_off=self.ball_off, ball_dist=self.ball_dist[:tot], ball_idx=self.ball_idx[:tot])

    def capture(self, queries: torch.Tensor):
        """Record the batched launches in a hipGraph (run eagerly first so the
        ball capacity is known)."""
        queries = queries.to(device=self.forest.device, dtype=self.forest.torch_dtype).contiguous()
        self.run(queries)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.launch(queries)
        torch.cuda.current_stream().wait_stream(s)
        self._graph = g
        return g

    def replay(self):
        """Replay the captured launches.  The graph writes the buffers the plan
        had at capture time; a ball total beyond their capacity (queries
        changed in place since the capture) raises, as does a replay after the
        plan was re-made (by a run that grew the capacity, or a config change):
        capture again then."""
        if self._graph is None:
            raise RuntimeError("no valid captured graph: call capture() (again) first")
        self._graph.replay()
        if self.has_ball:
            tot, ovf = self.status()
            if ovf:
                raise RuntimeError(f"ball results ({tot}) exceed the captured capacity ({self.capacity}): "
                                   "run() and capture() again")
        return self.results()

    def knn_rows(self, r: int):
        """(dist, idx) of set r as (nq_r, k_r) views."""
        k = int(self.k[r])
        nq = int(self.qoff[r + 1] - self.qoff[r])
        o = int(np.sum((self.qoff[1:r + 1] - self.qoff[:r]) * self.k[:r]))
        return self.knn_dist[o:o + nq * k].view(nq, k), self.knn_idx[o:o + nq * k].view(nq, k)
