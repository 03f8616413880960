"""The large single-GPU box at full scale.

N = 2^30 points in a periodic cube of side L = 500, float32 coordinates,
64-bit stored indices (the reference's LONGINT=1), 70 % of the points in
halos and 30 % uniform (the clustered recipe of ``synthetic``, generated on
the device).  Workload:

* the full GPU build (leaf 16);
* kNN, k = 16, of a random subset of 1e8 points (sorted rows);
* box search with 1e6 cubes of side log-uniform in [0.5, 5] centred
  uniformly in the box, so that many straddle the periodic faces.  The
  reference's box search does not wrap (cosmokdtree.f90:1260) and neither
  does this one: the part of a box outside [-L/2, L/2] holds no points.

Exactness at this scale is spot-checked on the device by brute force: the
kNN rows of a few queries against all N minimum-image distances (computed
in the reference's operation order, one rounding per operation), and the
box rows of a few boxes against the inclusive per-point test.

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
    ok_knn = 0
    for j in range(args.checks):
        r = j * (nq // args.checks)
        d2 = _min_image_d2(P, Q[r], L)
        ref_vals, _ = torch.topk(d2, k, largest=False, sorted=True)
        row_i, row_d = i[r], d[r]
        gd2 = d2[row_i]
        ok = torch.equal(torch.sort(gd2).values, ref_vals)
        ok = ok and torch.equal(torch.sqrt(gd2.double()).float(), row_d)
        ok = ok and len(set(row_i.tolist())) == k
        ok_knn += int(ok)
        del d2
    log(f"kNN spot checks: {ok_knn}/{args.checks} rows exact")
    del d, i, Q

    g.manual_seed(args.seed + 2)
    c = (torch.rand((nbox, 3), generator=g, device=dev, dtype=torch.float64) - 0.5) * L
    s = torch.exp(torch.empty(nbox, device=dev, dtype=torch.float64).uniform_(math.log(0.5), math.log(5.0),
                                                                                generator=g))
    lo, hi = c - s[:, None] / 2, c + s[:, None] / 2
    boxes = torch.stack([lo[:, 0], hi[:, 0], lo[:, 1], hi[:, 1], lo[:, 2], hi[:, 2]], dim=1).float().contiguous()
    straddle = float(((lo < -L / 2) | (hi > L / 2)).any(dim=1).float().mean())
    tree.box(boxes[:1000])  # warm-up
    torch.cuda.synchronize()
    e0.record()
    off, bi = tree.box(boxes)
    e1.record()
    torch.cuda.synchronize()
    box_ms = e0.elapsed_time(e1)
    total = int(off[-1])
    log(f"box search of {nbox} boxes ({straddle:.3f} straddle a face): {box_ms:.1f} ms, {total} results")
    ok_box = 0
    for j in range(args.checks):
        b = boxes[j * (nbox // args.checks)]
        m = ((P[:, 0] >= b[0]) & (P[:, 0] <= b[1]) & (P[:, 1] >= b[2]) & (P[:, 1] <= b[3]) &
             (P[:, 2] >= b[4]) & (P[:, 2] <= b[5]))
        want = set(torch.nonzero(m).flatten().tolist())
        r = j * (nbox // args.checks)
        got = set(bi[int(off[r]):int(off[r + 1])].tolist())
        ok_box += int(want == got)
    log(f"box spot checks: {ok_box}/{args.checks} rows exact")
    tree.free()
    print(json.dumps(dict(n=n, L=L, halo_frac=args.halo, idx_bits=int(info.idx_bits), nnodes=int(info.nnodes),
                          depth=int(info.depth), build_ms=build_ms, build_mpts_s=n / build_ms / 1e3,
                          knn_k=k, knn_queries=nq, knn_ms=knn_ms, knn_qps=nq / knn_ms * 1e3,
                          knn_checks_exact=f"{ok_knn}/{args.checks}", boxes=nbox, box_ms=box_ms,
                          boxes_per_s=nbox / box_ms * 1e3, box_results=total, box_straddle_frac=straddle,
                          box_checks_exact=f"{ok_box}/{args.checks}")), flush=True)


if __name__ == "__main__":
    main()
