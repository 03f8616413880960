"""Host-core timings of the reference and of the C restatement on the
benchmark configurations (test infrastructure: the checker side; nothing of
the product runs here).

    OMP_NUM_THREADS=8 OMP_PROC_BIND=close OMP_PLACES=cores \
        python -m oracle.calibrate --impl reference --case c1

Cases (seeded as the benchmarks are):
  c1      N = 1e6 uniform in [0,1)^3, float64, non-periodic, leaf 16; 1e5
          queries from the same distribution: build, kNN k=8 sorted, ball
          r=0.01, and box search with a cube of side 0.02 around each query,
          one box per call as the reference's Python API does
          (src/cosmokdtree/kdtree.py:82-87);
  c2_f32  the clustered periodic box, N = 2^24, L = 100, float32: build and
          kNN k=32 sorted of 2^20 of the points; ball r=0.2 of 1e6 of them;
  c2_f64  the same in float64.
One implementation per process (--impl reference | port): the reference's
OpenMP runtime binds the initial thread to one core, which would confine an
implementation timed after it.  Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from cosmokdtree_amd.synthetic import SEED_POINTS, SEED_QUERIES, SEED_SUBSET, clustered_periodic  # noqa: E402
from oracle.oracle import OracleTree, RefTree, ref_available  # noqa: E402


def timed(fn):
    t0 = time.perf_counter()
    r = fn()
    return time.perf_counter() - t0, r


def case_c1(cls, n_box_calls):
    rng = np.random.Generator(np.random.PCG64(SEED_POINTS))
    pts = rng.random((1_000_000, 3))
    q = np.random.Generator(np.random.PCG64(SEED_QUERIES)).random((100_000, 3))
    tb, tree = timed(lambda: cls(pts, 16, None))
    tk, _ = timed(lambda: tree.knn(q, 8, True))
    tr, _ = timed(lambda: tree.ball_csr(q, 0.01, False))
    boxes = np.empty((n_box_calls, 6))
    boxes[:, 0::2] = q[:n_box_calls] - 0.01
    boxes[:, 1::2] = q[:n_box_calls] + 0.01
    t0 = time.perf_counter()
    for b in boxes:  # one box per call, as kdtree.box_search does
        tree.box_csr(b)
    tx = time.perf_counter() - t0
    tree.close()
    return dict(build_mpts_s=1e6 / tb / 1e6, knn_k8_qps=1e5 / tk, ball_r001_qps=1e5 / tr,
                box_side002_calls_per_s=n_box_calls / tx, box_calls=n_box_calls)


def case_c2(cls, dtype, q_knn, q_ball):
    n = 1 << 24
    pts = clustered_periodic(n, 100.0, 0.5, seed=SEED_POINTS, dtype=dtype)
    rng = np.random.Generator(np.random.PCG64(SEED_QUERIES))
    q = pts[rng.choice(n, q_knn, replace=False)]
    sub = pts[np.random.Generator(np.random.PCG64(SEED_SUBSET)).choice(n, q_ball, replace=False)]
    tb, tree = timed(lambda: cls(pts, 16, np.full(3, 100.0)))
    tk, _ = timed(lambda: tree.knn(q, 32, True))
    tr, _ = timed(lambda: tree.ball_csr(sub, 0.2, False))
    tree.close()
    return dict(build_mpts_s=n / tb / 1e6, knn_k32_qps=q_knn / tk, knn_queries=q_knn, ball_r02_qps=q_ball / tr,
                ball_queries=q_ball)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--impl", choices=("reference", "port"), required=True)
    ap.add_argument("--case", choices=("c1", "c2_f32", "c2_f64"), required=True)
    ap.add_argument("--box-calls", type=int, default=100_000)
    ap.add_argument("--q-knn", type=int, default=1 << 20)
    ap.add_argument("--q-ball", type=int, default=1_000_000)
    args = ap.parse_args()
    dtype = np.float32 if args.case == "c2_f32" else np.float64
    periodic = args.case != "c1"
    if args.impl == "reference":
        if not ref_available(periodic, 3, dtype):
            print(json.dumps(dict(error="the reference libraries (oracle/_ref) are not built")))
            return
        cls = RefTree
    else:
        cls = OracleTree
    out = dict(impl=args.impl, case=args.case, threads=int(os.environ.get("OMP_NUM_THREADS", "0") or 0),
               affinity=len(os.sched_getaffinity(0)))
    if args.case == "c1":
        out.update(case_c1(cls, args.box_calls))
    else:
        out.update(case_c2(cls, dtype, args.q_knn, args.q_ball))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
