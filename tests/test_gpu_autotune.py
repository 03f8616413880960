"""The autotuner on a small batched workload: it stores its pick, the pick
is found again for the same workload, and the plan built with it returns
the results of the default plan (every configuration gives the same
results)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_tune_store_and_apply(ckd, tmp_path):
    from cosmokdtree_amd import autotune, batch
    from cosmokdtree_amd.synthetic import batched_sets
    path = str(tmp_path / "tune.json")
    pts, off, q, qoff, k, r = batched_sets(300000, 48, seed=5)
    nq = np.diff(qoff)
    print(f"\n[autotune] {len(pts)} points in {len(nq)} sets, {len(q)} queries", flush=True)
    best, rows = autotune.tune_batch(pts, off, q, qoff, k, r, [10.0] * 3, leaves=(8, 16), qpws=(16, 64), reps=1,
                                     path=path)
    assert best["ms"] == min(x["ms"] for x in rows if x["ms"] is not None)
    assert best["stack_depth"] in (32, 64) and best["stack_depth"] >= best["depth"]
    pick = autotune.lookup(best["key"], path)
    assert pick["leaf"] == best["leaf"] and pick["queries_per_wave"] == best["queries_per_wave"]
    P, Q = torch.from_numpy(pts).cuda(), torch.from_numpy(q).cuda()
    forest, plan, used = autotune.tuned_forest_and_plan(P, off, nq, k, r, [10.0] * 3, path=path)
    ref = batch.build_forest((P, off), leaf=16, boxsize=[10.0] * 3)
    try:
        assert used == pick and forest.leafsize == pick["leaf"]
        got = plan.run(Q)
        want = batch.BatchPlan(ref, nq, k, r).run(Q)
        assert torch.equal(got["knn_dist"], want["knn_dist"])
        assert torch.equal(got["ball_off"], want["ball_off"])
    finally:
        forest.free()
        ref.free()
