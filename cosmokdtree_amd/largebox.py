"""The large periodic box (2^30 points, L = 500) on one GPU at full scale.

N = 2^30 points in a periodic cube of side L = 500, float32 coordinates,
64-bit stored indices (the reference's LONGINT=1), 70 % of the points in
halos and 30 % uniform (the clustered recipe of ``synthetic``, generated on
the device).  Workload:

* the full GPU build (leaf 16);
* kNN, k = 16, of a random subset of 1e8 points (sorted rows): the plain
  search, and the learned-seed search (a network trained on an independent
  2^30-point draw of the same recipe), whose rows must equal the plain ones;
* the same queries through the batched engine (one-set forest of the same
  points, one plan with the kNN and a ball search of radius 0.05: the
  halos are dense enough that radius 0.5 returns ~5.6e10 rows);
* box search with 1e6 cubes of side log-uniform in [0.5, 5] centred
  uniformly in the box, taken modulo the box (``wrap``), so that the boxes
  that straddle a periodic face find the points beyond it.  The reference's
  box search does not wrap (cosmokdtree.f90:1260); this is the library's
  periodic rule: the union of the wrapped pieces, each searched with the
  reference's rules.

Exactness at this scale is spot-checked on the device by brute force: the
kNN rows of a few queries against all N minimum-image distances (computed
in the reference's operation order, one rounding per operation), the ball
rows against the inclusive per-point test, and the box rows against the
inclusive per-point test on the wrapped pieces.

    python -m cosmokdtree_amd.largebox [--n 1073741824] [--nq 1e8] [--nbox 1e6]

prints progress to stderr and one JSON line.  No reference counterpart
(the reference runs on the CPU and holds one tree per process).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from .synthetic import halo_sizes


def log(m):
    print(f"[largebox] {m}", file=sys.stderr, flush=True)


def clustered_periodic_device(n: int, L: float, halo_frac: float, seed: int, device="cuda",
                              chunk: int = 1 << 27) -> torch.Tensor:
    """The clustered periodic recipe (synthetic.clustered_periodic) drawn
    with torch on the device, float32, in chunks: uniform points first, then
    halos of dN/dn ~ n^-1.9 occupation around uniform centres with
    sigma = 0.02 n^(1/3), wrapped into [-L/2, L/2]."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    n_halo = int(round(n * halo_frac))
    n_uni = n - n_halo
    pts = torch.empty((n, 3), dtype=torch.float32, device=device)
    for a in range(0, n_uni, chunk):
        b = min(a + chunk, n_uni)
        pts[a:b] = (torch.rand((b - a, 3), generator=g, device=device) - 0.5) * L
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = halo_sizes(n_halo, rng)
    centres = torch.from_numpy((rng.random((len(sizes), 3)) - 0.5) * L).to(device=device, dtype=torch.float32)
    sig = torch.from_numpy(0.02 * np.cbrt(sizes.astype(np.float64))).to(device=device, dtype=torch.float32)
    first = torch.from_numpy(np.concatenate([[0], np.cumsum(sizes)[:-1]])).to(device)
    for a in range(0, n_halo, chunk):
        b = min(a + chunk, n_halo)
        pos = torch.arange(a, b, device=device)
        hid = torch.searchsorted(first, pos, right=True) - 1
        x = centres[hid] + sig[hid, None] * torch.randn((b - a, 3), generator=g, device=device)
        x = x - L * torch.floor((x + L / 2) / L)
        pts[n_uni + a:n_uni + b] = x.clamp_(-L / 2, L / 2)
        del pos, hid, x
    return pts


def _min_image_d2(P: torch.Tensor, q: torch.Tensor, L: float) -> torch.Tensor:
    """Squared minimum-image distances of all rows of P to q in the order of
    the reference's distance() (cosmokdtree.f90:1441-1465): dx = p - q,
    min(|dx|, L - |dx|), axis-order sum, one rounding per operation."""
    d2 = None
    for ax in range(3):
        a = (P[:, ax] - q[ax]).abs_()
        a = torch.minimum(a, L - a)
        t = a * a
        d2 = t if d2 is None else d2 + t
    return d2


def wrapped_box_mask(P: torch.Tensor, b: torch.Tensor, L: float) -> torch.Tensor:
    """Points inside box b (xmin, xmax, ymin, ...) taken modulo the periodic
    box: on every axis inside one of the pieces [lo - jL, hi - jL] clipped to
    [-L/2, L/2], j = m+1, m, m-1 with m = floor((lo + L/2) / L), the bounds
    rounded to float32 as the search rounds them; the whole axis when
    hi - lo >= L.  Inclusive, as the reference's leaf test."""
    H = np.float32(L / 2)
    inside = torch.ones(P.shape[0], dtype=torch.bool, device=P.device)
    for ax in range(3):
        lo, hi = np.float32(b[2 * ax].item()), np.float32(b[2 * ax + 1].item())
        if not lo <= hi:
            return torch.zeros_like(inside)
        if np.float32(hi - lo) >= np.float32(L):
            continue
        m = math.floor((float(lo) + float(H)) / float(L))
        x = P[:, ax]
        on = torch.zeros_like(inside)
        for j in (1, 0, -1):
            sh = np.float32(m + j) * np.float32(L)
            a_, b_ = max(np.float32(lo - sh), -H), min(np.float32(hi - sh), H)
            if a_ <= b_:
                on |= (x >= float(a_)) & (x <= float(b_))
        inside &= on
    return inside


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1 << 30)
    ap.add_argument("--L", type=float, default=500.0)
    ap.add_argument("--halo", type=float, default=0.7)
    ap.add_argument("--nq", type=float, default=1e8)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--nbox", type=float, default=1e6)
    ap.add_argument("--checks", type=int, default=8)
    ap.add_argument("--seed", type=int, default=20261205)
    ap.add_argument("--train-pairs", type=int, default=2_000_000)
    ap.add_argument("--ball-r", type=float, default=0.05)
    args = ap.parse_args()
    from . import kdtree

    dev = torch.device("cuda")
    n, L, k = args.n, args.L, args.k
    nq, nbox = int(args.nq), int(args.nbox)
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    t0 = time.time()
    P = clustered_periodic_device(n, L, args.halo, args.seed)
    torch.cuda.synchronize()
    log(f"generated {n} points on the device in {time.time() - t0:.1f}s")

    e0, e1 = ev(), ev()
    e0.record()
    tree = kdtree._build(P, 16, [L] * 3, idx_bits=64)
    e1.record()
    torch.cuda.synchronize()
    build_ms = e0.elapsed_time(e1)
    info = tree.info()
    log(f"build {build_ms:.1f} ms ({n / build_ms / 1e3:.1f} Mpts/s), nodes {info.nnodes}, depth {info.depth}, "
        f"idx_bits {info.idx_bits}, device memory {torch.cuda.max_memory_allocated() / 2**30:.1f} GiB (torch)")

    g = torch.Generator(device=dev)
    g.manual_seed(args.seed + 1)
    qi = torch.randint(0, n, (nq,), generator=g, device=dev)
    Q = P[qi]
    del qi
    tree.knn(Q[:1 << 20], k)  # warm-up
    torch.cuda.synchronize()
    e0.record()
    d, i = tree.knn(Q, k)
    e1.record()
    torch.cuda.synchronize()
    knn_ms = e0.elapsed_time(e1)
    log(f"kNN k={k} of {nq} queries: {knn_ms:.1f} ms ({nq / knn_ms / 1e3:.1f} Mq/s)")

    # kNN spot checks: the row's points are at exactly the k smallest squared
    # distances (so the row is a kNN set, ties aside), distinct, and the
    # reported distances are the correctly rounded square roots of theirs
    def knn_row_exact(row_d, row_i, q):
        d2 = _min_image_d2(P, q, L)
        ref_vals, _ = torch.topk(d2, k, largest=False, sorted=True)
        gd2 = d2[row_i]
        ok = torch.equal(torch.sort(gd2).values, ref_vals)
        ok = ok and torch.equal(torch.sqrt(gd2.double()).float(), row_d)
        return ok and len(set(row_i.tolist())) == k

    rows = [j * (nq // args.checks) for j in range(args.checks)]
    ok_knn = sum(int(knn_row_exact(d[r], i[r], Q[r])) for r in rows)
    log(f"kNN spot checks: {ok_knn}/{args.checks} rows exact")

    # learned seeds: a network trained on an independent draw of the recipe
    from . import learned
    t0 = time.time()
    T = clustered_periodic_device(n, L, args.halo, args.seed + 3)
    ttree = kdtree._build(T, 16, [L] * 3, idx_bits=64, keep_points=True)
    model = learned.fit(ttree, args.train_pairs, k=k, epochs=4, quantile=0.95, log=log)
    ttree.free()
    del T
    table = learned.feature_table(learned.occupancy(tree))
    torch.cuda.synchronize()
    train_s = time.time() - t0
    learned.seeded_knn(tree, model, Q[:1 << 20], k, table=table)  # warm-up
    torch.cuda.synchronize()
    e0.record()
    d_l, i_l = learned.seeded_knn(tree, model, Q, k, table=table)
    e1.record()
    torch.cuda.synchronize()
    seeded_ms = e0.elapsed_time(e1)
    seeded_equal = bool(torch.equal(d_l, d))
    del d_l, i_l
    st = torch.zeros(6, dtype=torch.int64, device=dev)
    learned.seeded_knn(tree, model, Q[:1 << 24], k, table=table, stats=st)  # fallback share (untimed)
    st = st.cpu().numpy()
    log(f"learned-seed kNN: {seeded_ms:.1f} ms ({nq / seeded_ms / 1e3:.1f} Mq/s, x{knn_ms / seeded_ms:.2f} vs plain), "
        f"distances equal to the plain rows: {seeded_equal}, fallback {st[2] / max(st[3], 1):.4f}, "
        f"training {train_s:.1f}s")

    # batched engine: the same queries, kNN and ball search in one plan
    from . import batch
    r_ball = args.ball_r
    forest = batch.build_forest((P, np.array([0, n])), leaf=16, boxsize=[L] * 3, idx_bits=64)
    plan = batch.BatchPlan(forest, [nq], [k], [r_ball], sorted=True)
    out = plan.run(Q)  # sizes the ball output
    torch.cuda.synchronize()
    e0.record()
    out = plan.run(Q)
    e1.record()
    torch.cuda.synchronize()
    batch_ms = e0.elapsed_time(e1)
    d_b, i_b = plan.knn_rows(0)
    batch_equal = bool(torch.equal(d_b, d))
    boff, bdist, bidx = out["ball_off"], out["ball_dist"], out["ball_idx"]
    ball_total = int(boff[-1])
    ok_ball = 0
    for r in rows:
        d2 = _min_image_d2(P, Q[r], L)
        want = set(torch.nonzero(d2 <= np.float32(r_ball) * np.float32(r_ball)).flatten().tolist())
        a, b = int(boff[r]), int(boff[r + 1])
        got_i = bidx[a:b]
        good = set(got_i.tolist()) == want and len(want) == b - a
        good = good and torch.equal(torch.sqrt(d2[got_i].double()).float(), bdist[a:b])
        ok_ball += int(good)
    log(f"batched plan (kNN k={k} + ball r={r_ball} of {nq} queries): {batch_ms:.1f} ms, kNN rows equal to the "
        f"single-tree rows: {batch_equal}, {ball_total} ball results, ball spot checks {ok_ball}/{args.checks} exact")
    del out, d_b, i_b, boff, bdist, bidx, plan
    forest.free()
    del d, i, Q

    g.manual_seed(args.seed + 2)
    c = (torch.rand((nbox, 3), generator=g, device=dev, dtype=torch.float64) - 0.5) * L
    s = torch.exp(torch.empty(nbox, device=dev, dtype=torch.float64).uniform_(math.log(0.5), math.log(5.0),
                                                                                generator=g))
    lo, hi = c - s[:, None] / 2, c + s[:, None] / 2
    boxes = torch.stack([lo[:, 0], hi[:, 0], lo[:, 1], hi[:, 1], lo[:, 2], hi[:, 2]], dim=1).float().contiguous()
    straddle = float(((lo < -L / 2) | (hi > L / 2)).any(dim=1).float().mean())
    tree.box(boxes[:1000], wrap=True)  # warm-up
    torch.cuda.synchronize()
    e0.record()
    off, bi = tree.box(boxes, wrap=True)
    e1.record()
    torch.cuda.synchronize()
    box_ms = e0.elapsed_time(e1)
    total = int(off[-1])
    log(f"box search of {nbox} boxes taken modulo the box ({straddle:.3f} straddle a face): {box_ms:.1f} ms, "
        f"{total} results")
    # spot checks: the boxes that straddle a face first
    strad_rows = torch.nonzero(((lo < -L / 2) | (hi > L / 2)).any(dim=1)).flatten()[:args.checks // 2].tolist()
    box_rows = strad_rows + [j * (nbox // args.checks) for j in range(args.checks - len(strad_rows))]
    ok_box = 0
    for r in box_rows:
        m = wrapped_box_mask(P, boxes[r], L)
        want = set(torch.nonzero(m).flatten().tolist())
        got = bi[int(off[r]):int(off[r + 1])].tolist()
        ok_box += int(want == set(got) and len(got) == len(want))
    log(f"box spot checks: {ok_box}/{len(box_rows)} rows exact ({len(strad_rows)} straddling)")
    tree.free()
    print(json.dumps(dict(n=n, L=L, halo_frac=args.halo, idx_bits=int(info.idx_bits), nnodes=int(info.nnodes),
                          depth=int(info.depth), build_ms=build_ms, build_mpts_s=n / build_ms / 1e3,
                          knn_k=k, knn_queries=nq, knn_ms=knn_ms, knn_qps=nq / knn_ms * 1e3,
                          knn_checks_exact=f"{ok_knn}/{args.checks}",
                          seeded_knn_ms=seeded_ms, seeded_speedup_vs_plain=knn_ms / seeded_ms,
                          seeded_equal_plain=seeded_equal, seeded_fallback=float(st[2] / max(st[3], 1)),
                          batch_ms=batch_ms, batch_ball_r=r_ball, batch_ball_results=ball_total,
                          batch_knn_equal_plain=batch_equal, batch_ball_checks_exact=f"{ok_ball}/{args.checks}",
                          boxes=nbox, box_wrap=True, box_ms=box_ms,
                          boxes_per_s=nbox / box_ms * 1e3, box_results=total, box_straddle_frac=straddle,
                          box_checks_exact=f"{ok_box}/{len(box_rows)}")), flush=True)


if __name__ == "__main__":
    main()
