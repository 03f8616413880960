#!/usr/bin/env python3
"""Benchmark of the k-d tree build and query path on one MI355X.

Workload (the clustered periodic box): N = 2^24 points in a periodic cube
of side L = 100 (coordinates in [-50, 50]), half uniform, half in halos
with dN/dn ~ n^-1.9 on [10, 1e5] and sigma = 0.02 n^(1/3), float32,
leaf size 16.  One step = a self-query kNN with k = 32 (sorted) of all N
points against the tree, inputs and outputs resident in HBM.  The tree is
built before the timed steps; its build rate is reported beside.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line (metric, value = queries/s, roofline of the kNN kernel,
host-core CPU baseline measured in the same run).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

# the host CPUs this process was given, read before torch loads: with
# OMP_PROC_BIND set in the environment, torch's OpenMP runtime binds the main
# thread to one CPU at its first parallel region, and children would inherit
# that mask.  The host-core baseline's children run on all of them.
AFFINITY0 = sorted(os.sched_getaffinity(0))

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "kNN k=32 queries/s on 16.7M periodic pts, learned-seed vs plain; build Mpts/s"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
PEAK_CLOCK_GHZ = 2.4  # MI355X peak engine clock


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def knn_bytes_per_query(D, s_c, s_i, k, n, nnodes, q):
    """Algorithmic HBM bytes per kNN query: query read, result write
    (distance in tree precision + int64 index) and the tree read once,
    amortised over the queries."""
    return D * s_c + k * (s_c + 8) + (n * (D * s_c + s_i) + nnodes * 16) / q


KNN_SOURCES = ("ckd_traverse.h", "ckd_query_impl.h", "ckd_internal.h", "ckd_knn_f32_per.hip")


def knn_source_sha():
    """Hash of the sources the kNN traversal kernel is compiled from."""
    import hashlib
    h = hashlib.sha256()
    for f in KNN_SOURCES:
        with open(os.path.join(ROOT, "cosmokdtree_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def build_bytes_per_point(D, s_c, s_i, n, leaf):
    """Algorithmic HBM bytes per point of the level-synchronous build: input
    read, read+write of coordinates and index per level, leaf-bucket write."""
    levels = int(np.ceil(np.log2(max(n / leaf, 2))))
    return D * s_c + 2 * (D * s_c + s_i) * levels + (D * s_c + s_i)


def cpu_baseline(n, q, k, cores):
    """Each implementation in a child process of its own (its own OpenMP
    runtime and CPU mask), OMP_* set before the runtime loads."""
    env = dict(os.environ)
    env.update(OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out = {}
    for impl in ("reference", "port"):
        cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--impl", impl, "--n", str(n), "--q", str(q),
               "--k", str(k)]
        p = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900,
                           preexec_fn=lambda: os.sched_setaffinity(0, AFFINITY0))
        if p.returncode != 0:
            log(f"cpu baseline ({impl}) failed: " + p.stderr[-2000:])
            continue
        r = json.loads(p.stdout.strip().splitlines()[-1])
        if impl in r:
            out[impl] = r[impl]
            out["affinity"] = r["affinity"]
            out["threads"] = r["threads"]
    return out or None


def knn_spot_checks(P, d, i, k, L, checks):
    """Rows of a full-size kNN result checked by brute force on the device:
    the row's points are distinct and sit at exactly the k smallest squared
    minimum-image distances of all N points (computed in the reference's
    operation order), and the reported distances are their correctly rounded
    square roots.  Returns the number of exact rows."""
    import torch

    from cosmokdtree_amd.largebox import _min_image_d2
    n = P.shape[0]
    ok = 0
    for j in range(checks):
        r = (j * 2654435761) % n  # spread over the box
        d2 = _min_image_d2(P, P[r], L)
        ref_vals = torch.topk(d2, k, largest=False, sorted=True).values
        row_i, row_d = i[r], d[r]
        gd2 = d2[row_i]
        good = torch.equal(torch.sort(gd2).values, ref_vals)
        good = good and torch.equal(torch.sqrt(gd2.double()).float(), row_d)
        good = good and len(set(row_i.tolist())) == k
        ok += int(good)
    return ok


def _ev_ms(fn, reps=3):
    """Median HIP-event time of fn() on the current stream (after one warm-up)."""
    import torch
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def calibration_table():
    """GPU rows of the calibration table (device-resident inputs and outputs):
    config 1 (1e6 uniform points in [0,1)^3, float64, 1e5 queries: build,
    kNN k=8 sorted, ball r=0.01, box side 0.02 for every query in one call)
    and the clustered periodic box (2^24 points, float32 and float64: build,
    self-kNN k=32 sorted, ball r=0.2 on 1e6 of the points)."""
    import torch

    from cosmokdtree_amd import kdtree
    from cosmokdtree_amd.synthetic import SEED_POINTS, SEED_QUERIES, SEED_SUBSET, clustered_periodic
    rows = {}
    rng = np.random.Generator(np.random.PCG64(SEED_POINTS))
    P = torch.from_numpy(rng.random((1_000_000, 3))).cuda()
    Q = torch.from_numpy(np.random.Generator(np.random.PCG64(SEED_QUERIES)).random((100_000, 3))).cuda()
    tb = _ev_ms(lambda: kdtree._build(P, 16, None).free())
    tree = kdtree._build(P, 16, None)
    boxes = torch.stack([Q[:, 0] - 0.01, Q[:, 0] + 0.01, Q[:, 1] - 0.01, Q[:, 1] + 0.01, Q[:, 2] - 0.01,
                         Q[:, 2] + 0.01], 1).contiguous()
    rows["c1_f64"] = dict(build_mpts_s=1e6 / tb / 1e3, knn_k8_qps=1e5 / _ev_ms(lambda: tree.knn(Q, 8)) * 1e3,
                          ball_r001_qps=1e5 / _ev_ms(lambda: tree.ball(Q, 0.01)) * 1e3,
                          box_side002_qps=1e5 / _ev_ms(lambda: tree.box(boxes)) * 1e3)
    tree.free()
    log(f"table c1: {rows['c1_f64']}")
    n = 1 << 24
    for tag, dt in (("c2_f32", np.float32), ("c2_f64", np.float64)):
        P = torch.from_numpy(clustered_periodic(n, 100.0, 0.5, seed=SEED_POINTS, dtype=dt)).cuda()
        box = np.full(3, 100.0)
        tb = _ev_ms(lambda: kdtree._build(P, 16, box).free())
        tree = kdtree._build(P, 16, box)
        sub = P[torch.from_numpy(np.random.Generator(np.random.PCG64(SEED_SUBSET)).choice(n, 1_000_000,
                                                                                      replace=False)).cuda()]
        rows[tag] = dict(build_mpts_s=n / tb / 1e3, knn_k32_qps=n / _ev_ms(lambda: tree.knn(P, 32)) * 1e3,
                         ball_r02_qps=1e6 / _ev_ms(lambda: tree.ball(sub, 0.2)) * 1e3)
        tree.free()
        del P, sub
        log(f"table {tag}: {rows[tag]}")
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host-core baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the float64 and ball-search extras")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-seed path")
    ap.add_argument("--train-pairs", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--quantile", type=float, default=0.95)
    ap.add_argument("--cpu-n", type=int, default=1 << 24)
    ap.add_argument("--cpu-q", type=int, default=1 << 20)
    ap.add_argument("--spot-checks", type=int, default=16,
                    help="rows of the last kNN step checked by brute force against all N points")
    ap.add_argument("--table", action="store_true",
                    help="print the GPU rows of the calibration table (configs 1 and 2) as one JSON line")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device copy of this size (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from cosmokdtree_amd import kdtree
    from cosmokdtree_amd.synthetic import SEED_POINTS, SEED_SUBSET, clustered_periodic

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.table:
        print(json.dumps(calibration_table()), flush=True)
        return
    dev = torch.device("cuda")
    N, K = args.n, args.k
    L = 100.0
    t0 = time.time()
    pts = clustered_periodic(N, L, 0.5, seed=SEED_POINTS, dtype=np.float32)
    log(f"generated {N} clustered periodic points in {time.time() - t0:.1f}s")
    P = torch.from_numpy(pts).to(dev)
    box = np.full(3, L)

    # ---- build (timed beside the steps)
    builds = []
    for i in range(args.warmup + 2):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        tree = kdtree._build(P, 16, box)
        e1.record()
        torch.cuda.synchronize()
        if i >= args.warmup:
            builds.append(e0.elapsed_time(e1))
        if i < args.warmup + 1:
            tree.free()
    build_ms = float(np.median(builds))
    info = tree.info()
    log(f"build {build_ms:.2f} ms ({N / build_ms / 1e3:.1f} Mpts/s), nodes={info.nnodes} depth={info.depth}")

    if args.calib_copy_mib:
        a = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        a.fill_(1)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    # ---- kNN steps
    for _ in range(args.warmup):
        d, i = tree.knn(P, K, sorted=True)
        del d, i
    torch.cuda.synchronize()
    stats = torch.zeros(6, dtype=torch.int64, device=dev)
    d, i = tree.knn(P, K, sorted=True, stats=stats)  # work counts of the whole step (untimed)
    st = stats.cpu().numpy()
    del d, i
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    tree.knn_timing(True)  # HIP events around each traversal kernel launch, on the launch stream
    e0.record()
    for _ in range(args.steps):
        d, i = tree.knn(P, K, sorted=True)
    e1.record()
    torch.cuda.synchronize()
    kms, kn = tree.knn_timing(False)
    plain_kernel_ms = kms / max(kn, 1)
    total_ms = e0.elapsed_time(e1)
    ms_step = total_ms / args.steps
    qps = N / (ms_step / 1e3)
    log(f"kNN k={K}: {ms_step:.2f} ms/step, {qps / 1e6:.1f} Mq/s")
    s_c, s_i = 4, (8 if info.idx_bits == 64 else 4)
    bq = knn_bytes_per_query(3, s_c, s_i, K, N, info.nnodes, N)
    plain_qps, plain_ms = qps, ms_step
    spot = knn_spot_checks(P, d, i, K, L, args.spot_checks)
    log(f"brute-force spot checks of the last step's rows: {spot}/{args.spot_checks} exact")
    d_plain = d
    del i

    # ---- learned seed: trained on an independent point set of the same recipe
    learned_info = None
    if not args.no_learned:
        from cosmokdtree_amd import learned
        from cosmokdtree_amd.synthetic import SEED_TRAIN
        t0 = time.time()
        tpts = torch.from_numpy(clustered_periodic(N, L, 0.5, seed=SEED_TRAIN, dtype=np.float32)).to(dev)
        ttree = kdtree._build(tpts, 16, box, keep_points=True)
        model = learned.fit(ttree, args.train_pairs, k=K, epochs=args.epochs, quantile=args.quantile, log=log)
        ttree.free()
        del tpts
        table = learned.feature_table(learned.occupancy(tree))
        train_s = time.time() - t0
        s1 = torch.zeros(6, dtype=torch.int64, device=dev)
        learned.seeded_knn(tree, model, P[:1 << 20], K, table=table, stats=s1)
        s1 = s1.cpu().numpy()
        for _ in range(args.warmup):
            d, i = learned.seeded_knn(tree, model, P, K, table=table)
            del d, i
        torch.cuda.synchronize()
        tree.knn_timing(True)
        e0.record()
        for _ in range(args.steps):
            d, i = learned.seeded_knn(tree, model, P, K, table=table)
        e1.record()
        torch.cuda.synchronize()
        kms, kn = tree.knn_timing(False)
        seeded_kernel_ms = kms / max(kn, 1)
        seeded_ms = e0.elapsed_time(e1) / args.steps
        exact = bool(torch.equal(d, d_plain))
        learned_info = dict(seeded_qps=N / (seeded_ms / 1e3), seeded_ms_per_step=seeded_ms,
                            speedup_vs_plain=plain_ms / seeded_ms, distances_identical=exact,
                            fallback_rate=float(s1[2] / max(s1[3], 1)),
                            nodes_visited_per_query=float(s1[0] / max(s1[3], 1)),
                            points_tested_per_query=float(s1[1] / max(s1[3], 1)),
                            insertions_per_query=float(s1[4] / max(s1[3], 1)), margin=model.margin,
                            traversal_kernel_ms=seeded_kernel_ms,
                            train_pairs=args.train_pairs, train_seconds=train_s)
        log(f"learned seed: {seeded_ms:.2f} ms/step ({N / seeded_ms / 1e3:.1f} Mq/s), "
            f"x{plain_ms / seeded_ms:.2f} vs plain, exact={exact}, fallback {learned_info['fallback_rate']:.4f}")
        del d, i
        # the network's forward pass alone over the step's features (bf16 MFMA
        # tiles): 2*(32*64 + 64*64 + 64) flops per query against the dense
        # bf16 peak
        X = learned.features(tree, P, table)
        fwd_ms = _ev_ms(lambda: model.seeds(X))
        del X
        flop_q = 2 * (32 * 64 + 64 * 64 + 64)
        learned_info["mlp_forward_ms"] = fwd_ms
        learned_info["mlp_forward_tflops"] = flop_q * N / (fwd_ms / 1e3) / 1e12
        learned_info["mlp_forward_mfma_frac"] = learned_info["mlp_forward_tflops"] / BF16_PEAK_TFLOPS
        log(f"MLP forward over {N} queries: {fwd_ms:.3f} ms, {learned_info['mlp_forward_tflops']:.1f} TFLOP/s "
            f"({100 * learned_info['mlp_forward_mfma_frac']:.2f} % of the bf16 dense peak)")
    del d_plain
    kernel_ms = plain_kernel_ms
    use_learned = bool(learned_info and learned_info["distances_identical"] and learned_info["seeded_qps"] > qps)
    if use_learned:
        qps, ms_step = learned_info["seeded_qps"], learned_info["seeded_ms_per_step"]
        kernel_ms = learned_info["traversal_kernel_ms"]
    # roofline of the dominant kernel: algorithmic bytes of one launch (all N
    # queries) over its average launch duration
    achieved = bq * N / (kernel_ms / 1e3) / 1e9
    sd = 32 if info.depth <= 32 else 64
    kernel_name = (f"k_knn_s<float, 3, true, {K}, {sd}, 1, false>" if info.depth <= 64
                   else f"k_knn_w<float, 3, true, {K}>")

    extra = dict(build_ms=build_ms, build_mpts_s=N / build_ms / 1e3, tree_nodes=int(info.nnodes),
                 knn_spot_checks_exact=f"{spot}/{args.spot_checks}",
                 tree_depth=int(info.depth), nodes_visited_per_query=float(st[0] / max(st[3], 1)),
                 points_tested_per_query=float(st[1] / max(st[3], 1)),
                 insertions_per_query=float(st[4] / max(st[3], 1)),
                 plain_qps=plain_qps, plain_ms_per_step=plain_ms, plain_traversal_kernel_ms=plain_kernel_ms,
                 learned=learned_info,
                 build_bytes_per_point=build_bytes_per_point(3, 4, s_i, N, 16))
    extra["build_roofline_frac"] = extra["build_bytes_per_point"] * N / (build_ms / 1e3) / 1e9 / HBM_PEAK_GBS

    if not args.no_extra:
        # SPH smoothing length + density of every point from its k neighbours,
        # fused into the traversal (writes 8 B per query instead of the k-row)
        h, rho = tree.sph_density(P, K)
        del h, rho
        torch.cuda.synchronize()
        tree.knn_timing(True)
        e0.record()
        for _ in range(args.steps):
            h, rho = tree.sph_density(P, K)
        e1.record()
        torch.cuda.synchronize()
        sms, sn = tree.knn_timing(False)
        sph_ms = e0.elapsed_time(e1) / args.steps
        extra["sph_density_qps"] = N / (sph_ms / 1e3)
        extra["sph_density_ms_per_step"] = sph_ms
        extra["sph_density_kernel_ms"] = sms / max(sn, 1)
        extra["sph_density_mean_rho"] = float(rho.double().mean().item())
        log(f"SPH density k={K}: {sph_ms:.2f} ms/step ({N / sph_ms / 1e3:.1f} Mq/s), "
            f"kernel {extra['sph_density_kernel_ms']:.2f} ms")
        del h, rho
        # ball search r = 0.2 on a random 1M subset (float32 tree)
        rng = np.random.Generator(np.random.PCG64(SEED_SUBSET))
        sub = torch.from_numpy(pts[rng.choice(N, 1 << 20, replace=False)]).to(dev)
        tree.ball(sub, 0.2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        off, bd, bi = tree.ball(sub, 0.2)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        extra["ball_r0.2_1M_qps"] = (1 << 20) / tb
        extra["ball_r0.2_mean_count"] = float(off[-1].item()) / (1 << 20)
        # algorithmic bytes per query: query read, offset, c_q rows of (f32
        # distance + int64 index); the whole call (one walk) over HBM peak
        bq_ball = 3 * 4 + 8 + extra["ball_r0.2_mean_count"] * (4 + 8)
        extra["ball_bytes_per_query"] = bq_ball
        extra["ball_roofline_frac"] = bq_ball * (1 << 20) / tb / 1e9 / HBM_PEAK_GBS
        log(f"ball r=0.2 on 1M: {tb * 1e3:.1f} ms, mean count {extra['ball_r0.2_mean_count']:.1f}")
        del off, bd, bi, sub
        # float64 build and kNN
        tree.free()
        P64 = P.double()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t64 = kdtree._build(P64, 16, box)
        torch.cuda.synchronize()
        extra["f64_build_mpts_s"] = N / (time.perf_counter() - t0) / 1e6
        d, i = t64.knn(P64, K, sorted=True)
        del d, i
        torch.cuda.synchronize()
        e0.record()
        d, i = t64.knn(P64, K, sorted=True)
        e1.record()
        torch.cuda.synchronize()
        extra["f64_knn_qps"] = N / (e0.elapsed_time(e1) / 1e3)
        log(f"f64: build {extra['f64_build_mpts_s']:.1f} Mpts/s, kNN {extra['f64_knn_qps'] / 1e6:.1f} Mq/s")
        del d, i, P64
        t64.free()
    else:
        tree.free()

    cpu = None
    if not args.no_cpu:
        # every core the process was given, unless OMP_NUM_THREADS sets the share
        aff = len(AFFINITY0)
        cores = min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or aff, aff)
        log(f"cpu baseline on {cores} host threads (affinity mask of {aff} CPUs at start, "
            f"{len(os.sched_getaffinity(0))} now) ...")
        res = cpu_baseline(args.cpu_n, args.cpu_q, K, cores)
        if res and ("reference" in res or "port" in res):
            main_impl = "reference" if "reference" in res else "port"
            r = res[main_impl]
            cpu = dict(value=r["knn_qps"], unit="queries/s", cores=cores, affinity=aff, kind=main_impl,
                       sample=f"clustered periodic recipe, n={args.cpu_n} float32 points, leaf 16: build, then "
                              f"kNN k={K} sorted of {args.cpu_q} of the points (OpenMP over queries), each "
                              f"implementation in its own process",
                       build_mpts_s=r["build_mpts_s"], build_s=r["build_s"], knn_s=r["knn_s"])
            if "port" in res and main_impl != "port":
                cpu["port_value"] = res["port"]["knn_qps"]
                cpu["port_build_mpts_s"] = res["port"]["build_mpts_s"]
            log(f"cpu: {json.dumps(cpu)}")

    # HBM bytes per launch from the committed rocprofv3 counter passes, used
    # only while the kernel sources are the ones they measured
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "knn_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("n") == N and tr.get("k") == K and tr.get("source_sha") == knn_source_sha():
            traffic = tr.get("seeded_bytes_per_launch" if use_learned else "plain_bytes_per_launch")
            if not use_learned and tr.get("plain_valu_insts"):
                # the kernel's other roofline: vector ALU issue.  A wave64
                # FP32 instruction without packed math holds a SIMD for 4
                # cycles (16 lanes a cycle); the counted instructions over
                # the SIMDs, against the launch at the peak engine clock
                simds = 4 * torch.cuda.get_device_properties(0).multi_processor_count
                cyc = tr["plain_valu_insts"] * 4 / simds
                extra["valu_issue"] = dict(valu_insts_per_launch=tr["plain_valu_insts"], simds=simds,
                                           cycles_per_simd=cyc, kernel_ms=kernel_ms, clock_ghz=PEAK_CLOCK_GHZ,
                                           frac=cyc / (kernel_ms * 1e-3 * PEAK_CLOCK_GHZ * 1e9),
                                           source="profiles/knn_clustered_2e24_k32_counters.txt")
                log(f"VALU issue of the traversal kernel: {cyc / 1e6:.1f} M cycles per SIMD in {kernel_ms:.2f} ms, "
                    f"{100 * extra['valu_issue']['frac']:.0f} % of the SIMDs' cycles at {PEAK_CLOCK_GHZ} GHz")
        else:
            log("profiles/knn_traffic.json measured other kernel sources or sizes: traffic not reported")

    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: clustered periodic box (seeded numpy PCG64), no dataset",
        "config": {"workload": "self-query kNN k=32 sorted, clustered periodic box N=2^24 L=100 halo 50%, "
                               "float32, leaf 16; value = the faster of the plain path and the learned-seed "
                               "path (features + MLP + seeded traversal, counted only when its distances "
                               "equal the plain path's)",
                   "n_points": N, "queries": N, "k": K, "leaf": 16, "boxsize": L,
                   "path": "learned-seed" if use_learned else "plain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel_name, "kernel_ms": kernel_ms, "bytes_per_query": bq},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    if spot != args.spot_checks:
        log(f"kNN rows differ from brute force ({spot}/{args.spot_checks} exact): the value is not valid")
        sys.exit(3)


if __name__ == "__main__":
    main()
