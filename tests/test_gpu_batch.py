"""GPU tests of the forest build and the batched persistent query engine.

Each set's batched results must equal the oracle's results for that set
alone (distances bit-exact, ties at the k-th distance as sets, ball rows as
sets); a captured hipGraph must reproduce the eager results on every replay,
with eager work and synchronisations between the replays.
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleTree
from tests.parity import assert_csr_equal, assert_knn_rows

pytestmark = pytest.mark.gpu

L = 10.0


@pytest.fixture(scope="module")
def workload(ckd):
    from cosmokdtree_amd.synthetic import halo_set
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(20), np.log(20000), 24)).astype(int)
    sizes[:3] = [1, 5, 16]  # single point, tiny, exactly one leaf
    sets = [halo_set(int(n), L, seed=100 + i) for i, n in enumerate(sizes)]
    nq = [max(1, int(n) // 10) for n in sizes]
    ks = rng.choice([8, 16, 32, 64], len(sizes)).astype(np.int32)
    radii = np.exp(rng.uniform(np.log(0.05), np.log(0.5), len(sizes)))
    qs = [s[rng.integers(0, len(s), m)] for s, m in zip(sets, nq)]
    return sets, nq, ks, radii, qs


@pytest.mark.parametrize("kernel,qpw,leaf", [("stack", 64, 16), ("stack", 16, 16), ("stack", 4, 8), ("stack", 1, 32),
                                              ("lane", 64, 16), ("lane", 64, 64)])
def test_batch_matches_per_set_oracle(ckd, workload, kernel, qpw, leaf):
    from cosmokdtree_amd import batch
    sets, nq, ks, radii, qs = workload
    forest = batch.build_forest(sets, leaf=leaf, boxsize=[L] * 3)
    assert forest.info().nsets == len(sets)
    plan = batch.BatchPlan(forest, nq, ks, radii, sorted=True, threads=128, knn_kernel=kernel,
                           queries_per_wave=qpw)
    out = plan.run(torch.from_numpy(np.concatenate(qs)).cuda())
    off = out["ball_off"].cpu().numpy()
    bd, bi = out["ball_dist"].cpu().numpy(), out["ball_idx"].cpu().numpy()
    for r, (pts, q) in enumerate(zip(sets, qs)):
        orc = OracleTree(pts, leaf, np.full(3, L))
        k = int(ks[r])
        d_ref, i_ref = orc.knn(q, k, True)
        d_got, i_got = plan.knn_rows(r)
        assert_knn_rows(pts, q, k, d_ref, i_ref, d_got.cpu().numpy(), i_got.cpu().numpy(), L=np.full(3, L))
        o_ref, bd_ref, bi_ref = orc.ball_csr(q, radii[r], False)
        q0 = int(plan.qoff[r])
        rows = off[q0:q0 + len(q) + 1]
        assert_csr_equal(o_ref, bd_ref, bi_ref, rows - rows[0], bd[rows[0]:rows[-1]], bi[rows[0]:rows[-1]])
    forest.free()


@pytest.mark.parametrize("periodic,dtype", [(False, np.float32), (True, np.float64), (False, np.float64)])
@pytest.mark.parametrize("kernel", ["stack", "lane"])
def test_batch_open_and_f64_forests(ckd, workload, periodic, dtype, kernel):
    """The other compiled variants of the engine (non-periodic forests,
    float64 coordinates): the same per-set comparison with the oracle."""
    from cosmokdtree_amd import batch
    sets, nq, ks, radii, qs = workload
    sets = [s.astype(dtype) for s in sets]
    qs = [q.astype(dtype) for q in qs]
    box = np.full(3, L) if periodic else None
    forest = batch.build_forest(sets, leaf=16, boxsize=None if box is None else list(box))
    try:
        plan = batch.BatchPlan(forest, nq, ks, radii, sorted=True, threads=128, knn_kernel=kernel,
                               queries_per_wave=64)
        out = plan.run(torch.from_numpy(np.concatenate(qs)).cuda())
        off = out["ball_off"].cpu().numpy()
        bd, bi = out["ball_dist"].cpu().numpy(), out["ball_idx"].cpu().numpy()
        assert bd.dtype == dtype
        for r, (pts, q) in enumerate(zip(sets, qs)):
            orc = OracleTree(pts, 16, box)
            k = int(ks[r])
            d_ref, i_ref = orc.knn(q, k, True)
            d_got, i_got = plan.knn_rows(r)
            assert_knn_rows(pts, q, k, d_ref, i_ref, d_got.cpu().numpy(), i_got.cpu().numpy(), L=box)
            o_ref, bd_ref, bi_ref = orc.ball_csr(q, radii[r], False)
            q0 = int(plan.qoff[r])
            rows = off[q0:q0 + len(q) + 1]
            assert_csr_equal(o_ref, bd_ref, bi_ref, rows - rows[0], bd[rows[0]:rows[-1]], bi[rows[0]:rows[-1]])
            orc.close()
    finally:
        forest.free()


def test_batch_sets_without_ball(ckd, workload):
    """Sets with radius 0 (and k = 0) take part without a ball search: their
    ball rows are empty and the other sets' rows still match the oracle."""
    from cosmokdtree_amd import batch
    sets, nq, ks, radii, qs = workload
    radii = radii.copy()
    radii[::3] = 0.0
    ks = ks.copy()
    ks[1::5] = 0
    forest = batch.build_forest(sets, leaf=16, boxsize=[L] * 3)
    try:
        plan = batch.BatchPlan(forest, nq, ks, radii, sorted=True, threads=128, knn_kernel="stack",
                               queries_per_wave=64)
        for rep in range(2):  # the second run reuses the plan and its buffers
            out = plan.run(torch.from_numpy(np.concatenate(qs)).cuda())
            off = out["ball_off"].cpu().numpy()
            bd, bi = out["ball_dist"].cpu().numpy(), out["ball_idx"].cpu().numpy()
            for r, (pts, q) in enumerate(zip(sets, qs)):
                q0 = int(plan.qoff[r])
                rows = off[q0:q0 + len(q) + 1]
                if radii[r] == 0:
                    assert rows[-1] == rows[0], f"set {r} has no ball search but {rows[-1] - rows[0]} rows"
                    continue
                orc = OracleTree(pts, 16, np.full(3, L))
                o_ref, bd_ref, bi_ref = orc.ball_csr(q, radii[r], False)
                assert_csr_equal(o_ref, bd_ref, bi_ref, rows - rows[0], bd[rows[0]:rows[-1]], bi[rows[0]:rows[-1]])
                if ks[r] > 0:
                    d_ref, i_ref = orc.knn(q, int(ks[r]), True)
                    d_got, i_got = plan.knn_rows(r)
                    assert_knn_rows(pts, q, int(ks[r]), d_ref, i_ref, d_got.cpu().numpy(), i_got.cpu().numpy(),
                                    L=np.full(3, L))
                orc.close()
        del plan
    finally:
        forest.free()


def test_batch_graph_replays(ckd, workload):
    from cosmokdtree_amd import batch
    sets, nq, ks, radii, qs = workload
    forest = batch.build_forest(sets, leaf=16, boxsize=[L] * 3)
    plan = batch.BatchPlan(forest, nq, ks, radii, sorted=True)
    q = torch.from_numpy(np.concatenate(qs)).cuda()
    eager = {k: v.clone() for k, v in plan.run(q).items()}
    plan.capture(q)
    for rep in range(3):
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        got = plan.replay()
        torch.cuda.synchronize()
        for key in ("knn_dist", "knn_idx", "ball_off"):
            assert torch.equal(got[key], eager[key]), (rep, key)
        # ball rows are written in traversal order: same per-row sets each replay
        assert torch.equal(torch.sort(got["ball_idx"])[0], torch.sort(eager["ball_idx"])[0])
        del junk
    forest.free()


def test_batch_replan_invalidates_graph(ckd, workload):
    """A run that outgrows the ball capacity re-makes the plan: the graph
    captured on the old buffers must not replay, and a plan whose sets all
    have radius 0 reports empty ball rows."""
    from cosmokdtree_amd import batch
    sets, nq, ks, radii, qs = workload
    forest = batch.build_forest(sets, leaf=16, boxsize=[L] * 3)
    try:
        q = torch.from_numpy(np.concatenate(qs)).cuda()
        plan = batch.BatchPlan(forest, nq, ks, radii * 0.5, sorted=True)
        plan.capture(q)
        small_cap = plan.capacity
        plan.radius = np.ascontiguousarray(radii * 1.0)
        plan._make(small_cap)  # same capacity, larger radii: the next run overflows and re-plans
        with pytest.raises(RuntimeError):
            plan.replay()
        out = plan.run(q)
        assert plan.capacity > small_cap and plan._graph is None
        assert int(out["ball_off"][-1]) == plan.capacity
        with pytest.raises(RuntimeError):
            plan.replay()
        none = batch.BatchPlan(forest, nq, ks, np.zeros(len(nq)), sorted=True)
        out = none.run(q)
        assert not none.has_ball and torch.count_nonzero(out["ball_off"]) == 0 and out["ball_idx"].numel() == 0
    finally:
        forest.free()


@pytest.mark.parametrize("periodic,dtype,threads", [(True, np.float32, 256), (False, np.float64, 128),
                                                    (True, np.float32, 64)])
def test_batch_shared_ball_walk(ckd, workload, periodic, dtype, threads):
    """Ball items on the shared walk (one walk per wave over each set's wide
    records): every set's rows equal the oracle's, also when the first run
    overflows the ball capacity and the plan is re-made."""
    from cosmokdtree_amd import batch
    sets, nq, ks, radii, qs = workload
    sets = [s.astype(dtype) for s in sets]
    qs = [q.astype(dtype) for q in qs]
    box = [L] * 3 if periodic else None
    forest = batch.build_forest(sets, leaf=16, boxsize=box)
    print(f"\n[shared ball walk] periodic={periodic} {np.dtype(dtype).name} threads={threads}", flush=True)
    try:
        plan = batch.BatchPlan(forest, nq, ks, radii, sorted=True, threads=threads, ball_kernel="stack",
                               ball_capacity=64)
        out = plan.run(torch.from_numpy(np.concatenate(qs)).cuda())
        assert plan.capacity > 64  # the first run overflowed and the plan was re-made
        off = out["ball_off"].cpu().numpy()
        bd, bi = out["ball_dist"].cpu().numpy(), out["ball_idx"].cpu().numpy()
        lane = batch.BatchPlan(forest, nq, ks, radii, sorted=True, threads=threads, ball_kernel="lane")
        lo = lane.run(torch.from_numpy(np.concatenate(qs)).cuda())
        assert torch.equal(lo["ball_off"].cpu(), torch.from_numpy(off))
        for r, (pts, q) in enumerate(zip(sets, qs)):
            orc = OracleTree(pts, 16, None if box is None else np.full(3, L))
            o_ref, bd_ref, bi_ref = orc.ball_csr(q, radii[r], False)
            q0 = int(plan.qoff[r])
            rows = off[q0:q0 + len(q) + 1]
            assert_csr_equal(o_ref, bd_ref, bi_ref, rows - rows[0], bd[rows[0]:rows[-1]], bi[rows[0]:rows[-1]])
            orc.close()
    finally:
        forest.free()
