"""Autotuner of the batched engine.

Sweeps, for one batched workload (many point sets, each with its own query
batch, k and ball radius):

* leaf size {8, 16, 32, 64}: the forest is rebuilt for each;
* queries per wave {1, 4, 16, 64} of the stack kNN walk (how many queries
  share one wave; each query is held by 64 / that many lanes), plus the
  one-walk-per-query kernel;
* stack depth {32, 64} of the stack walk (a depth below the forest's depth
  is skipped: it is never used);

timing ``ckd_batch_run`` launches with HIP events -- the kNN work (leaf
ordering included) per configuration, the ball work (count, scan, fill) once
per leaf size since it does not depend on the kNN settings -- and returns
the fastest configuration with the full table.  Every configuration returns the same
results, so the choice only changes speed.

    python -m cosmokdtree_amd.autotune [--total 200000000] [--sets 4096]

runs it on the batched workload of ``synthetic.batched_sets`` and prints the
table and one JSON line.  No reference counterpart.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

LEAVES = (8, 16, 32, 64)
QPWS = (1, 4, 16, 64)
STACKS = (32, 64)


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
               lane=True, reps=3, threads=256, log=None):
    """Sweep the configurations; returns (best dict, list of result dicts).
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
        # the ball phase does not depend on the kNN settings: timed once per leaf
        ball_ms = None
        if radius is not None:
            bplan = batch.BatchPlan(forest, nq, np.zeros(len(nq), np.int32), radius, threads=threads)
            ball_ms = _time_plan(bplan, q, reps)
            del bplan
        plan = batch.BatchPlan(forest, nq, k, None, sorted=True, threads=threads)
        configs = [("stack", w, s) for w in qpws for s in stacks] + ([("lane", 64, 0)] if lane else [])
        for kernel, qpw, sd in configs:
            row = dict(leaf=leaf, kernel=kernel, queries_per_wave=qpw, stack_depth=sd, depth=int(info.depth),
                       build_ms=float(info.build_ms), ball_ms=ball_ms)
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
    return best, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--total", type=float, default=2e8)
    ap.add_argument("--sets", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    from .synthetic import batched_sets

    def log(m):
        print(f"[autotune] {m}", file=sys.stderr, flush=True)

    t0 = time.time()
    pts, off, q, qoff, k, r = batched_sets(int(args.total), args.sets)
    log(f"generated {len(pts)} points in {args.sets} sets, {len(q)} queries in {time.time() - t0:.1f}s")
    best, rows = tune_batch(pts, off, q, qoff, k, r, [10.0] * 3, reps=args.reps, log=log)
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
