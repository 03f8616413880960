"""Autotuner of the batched engine.

Sweeps, for one batched workload (many point sets, each with its own query
batch, k and ball radius):

* leaf size {8, 16, 32, 64}: the forest is rebuilt for each;
* queries per wave {1, 4, 16, 64} of the stack kNN walk (how many queries
  share one wave; each query is held by 64 / that many lanes), plus the
  one-walk-per-query kNN kernel;
* the ball kernel: one walk per query, or one shared walk per wave;
* stack depth of the stack walk: the smallest of {32, 64} that holds the
  forest's depth.  The stack only sizes an LDS array; a deeper one than the
  tree needs was measured equal within 1 % on the 2e8-point workload, so it
  is set by that rule and swept only with ``stacks=(32, 64)`` /
  ``--sweep-stack`` (a depth below the forest's depth is never used);

timing ``ckd_batch_run`` launches with HIP events -- the kNN work (leaf
ordering included) per configuration, the ball work (count, scan, fill) once
per leaf size since it does not depend on the kNN settings -- and returns
the fastest configuration with the full table.  Every configuration returns
the same results, so the choice only changes speed.

The pick is stored per (device, workload signature) in a JSON cache
(``$CKD_AUTOTUNE_CACHE``, default ``~/.cache/cosmokdtree_amd/autotune.json``):
``lookup`` finds it again and ``tuned_forest_and_plan`` builds the forest and
plan with it (defaults when the workload was never tuned).  The signature
rounds the workload's sizes to powers of two and its k mix to tenths, so
workloads of the same shape share a pick.

    python -m cosmokdtree_amd.autotune [--total 200000000] [--sets 4096]

runs it on the batched workload of ``synthetic.batched_sets``, stores the pick
and prints the table and one JSON line.  No reference counterpart.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

LEAVES = (8, 16, 32, 64)
QPWS = (1, 4, 16, 64)
STACKS = (0,)  # 0: the smallest stack that holds the forest's depth


def cache_path() -> str:
    return os.environ.get("CKD_AUTOTUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache",
                                                                 "cosmokdtree_amd", "autotune.json")


def device_name() -> str:
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}:{getattr(p, 'gcnArchName', '')}:{p.multi_processor_count}cu"


def workload_key(set_sizes, nq_per_set, k_per_set, radius_per_set=None, dtype="float32", periodic=True,
                 device: str | None = None) -> str:
    """Signature of a batched workload: device, precision, periodicity, and
    the sizes rounded to powers of two (points, sets, queries, largest set),
    the share of sets per k (tenths) and whether a ball search runs."""
    sizes = np.asarray(set_sizes, dtype=np.int64)
    nq = np.asarray(nq_per_set, dtype=np.int64)
    ks = np.asarray(k_per_set, dtype=np.int64)
    p2 = lambda x: int(round(math.log2(max(int(x), 1))))  # noqa: E731
    mix = ",".join(f"{int(k)}:{round(float(np.mean(ks == k)), 1)}" for k in sorted(set(ks.tolist())))
    ball = radius_per_set is not None and bool(np.any(np.asarray(radius_per_set) > 0))
    dev = device if device is not None else device_name()
    return (f"{dev}|{np.dtype(dtype).name}|{'per' if periodic else 'open'}|n2^{p2(sizes.sum())}|sets2^{p2(len(sizes))}"
            f"|q2^{p2(nq.sum())}|max2^{p2(sizes.max() if len(sizes) else 1)}|k{mix}|ball{int(ball)}")


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(key: str, best: dict, path: str | None = None) -> None:
    """Store a pick (a row of tune_batch) under `key`."""
    path = path or cache_path()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    data = _load_cache(path)
    data[key] = {k: best[k] for k in ("leaf", "kernel", "queries_per_wave", "stack_depth", "ball_kernel", "ms",
                                      "knn_ms", "ball_ms") if k in best}
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def lookup(key: str, path: str | None = None) -> dict | None:
    """The stored pick for `key`, or None."""
    return _load_cache(path or cache_path()).get(key)


def tuned_forest_and_plan(points, offsets, nq_per_set, k_per_set, radius_per_set=None, boxsize=None, sorted=True,
                          threads=256, path: str | None = None):
    """Forest and BatchPlan of a workload with its stored pick applied (leaf
    size, kNN kernel, queries per wave, stack depth); the library's defaults
    when it was never tuned.  Returns (forest, plan, pick or None)."""
    from . import batch
    pts = torch.as_tensor(points)
    off = np.asarray(offsets, dtype=np.int64)
    key = workload_key(np.diff(off), nq_per_set, k_per_set, radius_per_set,
                       "float32" if pts.dtype == torch.float32 else "float64", boxsize is not None)
    pick = lookup(key, path)
    leaf = int(pick["leaf"]) if pick else 16
    forest = batch.build_forest((pts, off), leaf=leaf, boxsize=boxsize)
    if pick:
        plan = batch.BatchPlan(forest, nq_per_set, k_per_set, radius_per_set, sorted=sorted, threads=threads,
                               knn_kernel=pick["kernel"], queries_per_wave=int(pick["queries_per_wave"]),
                               stack_depth=int(pick["stack_depth"]), ball_kernel=pick.get("ball_kernel", "lane"))
    else:
        plan = batch.BatchPlan(forest, nq_per_set, k_per_set, radius_per_set, sorted=sorted, threads=threads)
    return forest, plan, pick


def _time_plan(plan, q, reps):
    plan.run(q)  # warm-up; also sizes the ball output
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.launch(q)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    return float(np.median(times))


def tune_batch(points, offsets, queries, qoffsets, k, radius, boxsize, leaves=LEAVES, qpws=QPWS, stacks=STACKS,
               lane=True, reps=3, threads=256, log=None, store: bool = True, path: str | None = None):
    """Sweep the configurations; returns (best dict, list of result dicts)
    and, with store, keeps the best under the workload's key (``lookup``).
    points (N, 3) and queries (Q, 3) may be host arrays or device tensors."""
    from . import batch
    pts = torch.as_tensor(points).cuda()
    q = torch.as_tensor(queries).to(device="cuda", dtype=pts.dtype).contiguous()
    nq = np.diff(np.asarray(qoffsets, dtype=np.int64))
    results = []
    for leaf in leaves:
        t0 = time.time()
        forest = batch.build_forest((pts, np.asarray(offsets, dtype=np.int64)), leaf=leaf, boxsize=boxsize)
        info = forest.info()
        if log:
            log(f"leaf {leaf}: forest of {info.nsets} sets built in {info.build_ms:.1f} ms, depth {info.depth}")
        # the ball phase does not depend on the kNN settings: timed once per
        # leaf size, for both ball kernels (the faster is kept)
        ball_ms, ball_kernel = None, "lane"
        if radius is not None:
            for bk in ("lane", "stack"):
                bplan = batch.BatchPlan(forest, nq, np.zeros(len(nq), np.int32), radius, threads=threads,
                                        ball_kernel=bk)
                ms = _time_plan(bplan, q, reps)
                if log:
                    log(f"leaf {leaf}: ball work {ms:.2f} ms with the {bk} kernel")
                if ball_ms is None or ms < ball_ms:
                    ball_ms, ball_kernel = ms, bk
                del bplan
        plan = batch.BatchPlan(forest, nq, k, None, sorted=True, threads=threads)
        configs = [("stack", w, s) for w in qpws for s in stacks] + ([("lane", 64, 0)] if lane else [])
        for kernel, qpw, sd in configs:
            if sd == 0:
                sd = 32 if info.depth <= 32 else 64
            row = dict(leaf=leaf, kernel=kernel, queries_per_wave=qpw, stack_depth=sd, depth=int(info.depth),
                       build_ms=float(info.build_ms), ball_ms=ball_ms, ball_kernel=ball_kernel)
            if kernel == "stack" and sd < info.depth:
                row["knn_ms"] = row["ms"] = None
                row["skipped"] = "stack shorter than the tree"
            else:
                plan.config(kernel, qpw, sd)
                row["knn_ms"] = _time_plan(plan, q, reps)
                row["ms"] = row["knn_ms"] + (ball_ms or 0.0)
            results.append(row)
            if log:
                log(json.dumps(row))
        del plan
        forest.free()
        if log:
            log(f"leaf {leaf} done in {time.time() - t0:.1f}s")
    timed = [r for r in results if r["ms"] is not None]
    best = min(timed, key=lambda r: r["ms"])
    if store:
        key = workload_key(np.diff(np.asarray(offsets, dtype=np.int64)), nq, k, radius,
                           "float32" if pts.dtype == torch.float32 else "float64", boxsize is not None)
        save(key, best, path)
        best = dict(best, key=key)
    return best, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--total", type=float, default=2e8)
    ap.add_argument("--sets", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--sweep-stack", action="store_true", help="also sweep stack depth 32 / 64")
    args = ap.parse_args()
    from .synthetic import batched_sets

    def log(m):
        print(f"[autotune] {m}", file=sys.stderr, flush=True)

    t0 = time.time()
    pts, off, q, qoff, k, r = batched_sets(int(args.total), args.sets)
    log(f"generated {len(pts)} points in {args.sets} sets, {len(q)} queries in {time.time() - t0:.1f}s")
    best, rows = tune_batch(pts, off, q, qoff, k, r, [10.0] * 3, reps=args.reps, log=log,
                            stacks=(32, 64) if args.sweep_stack else STACKS)
    log(f"stored the pick under {best['key']} in {cache_path()}")
    nqt = int(qoff[-1])
    print(f"{'leaf':>4} {'kernel':>6} {'q/wave':>6} {'stack':>5} {'kNN ms':>9} {'ball ms':>9} {'total':>9}  "
          f"queries/s", file=sys.stderr)
    for row in rows:
        ms = row["ms"]
        if ms is None:
            print(f"{row['leaf']:>4} {row['kernel']:>6} {row['queries_per_wave']:>6} {row['stack_depth']:>5}   skipped",
                  file=sys.stderr)
            continue
        print(f"{row['leaf']:>4} {row['kernel']:>6} {row['queries_per_wave']:>6} {row['stack_depth']:>5} "
              f"{row['knn_ms']:9.2f} {row['ball_ms'] or 0:9.2f} {ms:9.2f}  {nqt / ms * 1e3:.3e}", file=sys.stderr)
    print(json.dumps(dict(workload=dict(points=len(pts), sets=args.sets, queries=nqt,
                                        k_sets={int(x): int((k == x).sum()) for x in (8, 16, 32, 64)}),
                          best=best, best_queries_per_s=nqt / best["ms"] * 1e3, table=rows)), flush=True)


if __name__ == "__main__":
    main()
