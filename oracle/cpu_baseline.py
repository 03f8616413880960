"""Host-core CPU baseline for bench.py (test infrastructure: the checker side).

Run as a child process, one per implementation, so the OpenMP variables
are in place before the OpenMP runtime loads and each implementation has
the process to itself:

    OMP_NUM_THREADS=16 OMP_PROC_BIND=close OMP_PLACES=cores \
        python -m oracle.cpu_baseline --impl reference --n 16777216 --q 1048576 --k 32

(The reference's LLVM OpenMP runtime binds the process's initial thread to
one core under OMP_PROC_BIND=close; an implementation timed after it in the
same process would run its whole team inside that one-core mask.)

Times, on the clustered periodic workload (float32, L=100, halo fraction
0.5; by default all 2^24 points and 2^20 of them as queries):
  * --impl reference: oracle/_ref, compiled from src/cosmokdtree.f90: build,
    then kNN with an OpenMP parallel loop over queries as its f2py bridge
    does (cosmokdtree2py.f90:132-141);
  * --impl port: the C restatement (oracle/libckd_oracle.so), same steps.
Prints one JSON object with the thread count and the size of the process's
CPU affinity mask.
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

from cosmokdtree_amd.synthetic import SEED_POINTS, SEED_QUERIES, clustered_periodic  # noqa: E402
from oracle.oracle import OracleTree, RefTree, ref_available  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--impl", choices=("reference", "port"), default="reference")
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--q", type=int, default=1 << 20)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--L", type=float, default=100.0)
    ap.add_argument("--dtype", default="f32")
    args = ap.parse_args()
    dt = np.float32 if args.dtype == "f32" else np.float64
    pts = clustered_periodic(args.n, args.L, 0.5, seed=SEED_POINTS, dtype=dt)
    rng = np.random.Generator(np.random.PCG64(SEED_QUERIES))
    q = pts[rng.choice(args.n, args.q, replace=False)]
    L = np.full(3, args.L)
    out = dict(impl=args.impl, n=args.n, q=args.q, k=args.k,
               threads=int(os.environ.get("OMP_NUM_THREADS", "0") or 0), affinity=len(os.sched_getaffinity(0)),
               dtype=args.dtype)
    impls = [("port", OracleTree)] if args.impl == "port" else []
    if args.impl == "reference":
        if not ref_available(True, 3, dt):
            print(json.dumps(dict(out, error="the reference libraries (oracle/_ref) are not built")))
            return
        impls = [("reference", RefTree)]
    for name, cls in impls:
        t0 = time.perf_counter()
        tree = cls(pts, 16, L)
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        tree.knn(q, args.k, True)
        tq = time.perf_counter() - t0
        tree.close()
        out[name] = dict(build_s=tb, build_mpts_s=args.n / tb / 1e6, knn_s=tq, knn_qps=args.q / tq)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
