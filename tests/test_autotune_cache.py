"""The autotuner's cache of picks (no GPU): workload signatures and the
JSON store round trip."""
import numpy as np

from cosmokdtree_amd import autotune


def test_workload_key_rounds_shapes():
    sizes = np.array([1000, 20000, 500000])
    k = np.array([8, 32, 32])
    a = autotune.workload_key(sizes, sizes // 10, k, [0.1, 0.2, 0.3], "float32", True, device="gfx950")
    b = autotune.workload_key(sizes * 1.1, sizes // 10, k, [0.2, 0.1, 0.3], "float32", True, device="gfx950")
    assert a == b  # same shape: sizes within a power of two, same k mix
    assert a != autotune.workload_key(sizes * 4, sizes // 10, k, [0.1] * 3, "float32", True, device="gfx950")
    assert a != autotune.workload_key(sizes, sizes // 10, k, None, "float32", True, device="gfx950")
    assert a != autotune.workload_key(sizes, sizes // 10, k, [0.1] * 3, "float64", True, device="gfx950")
    assert a != autotune.workload_key(sizes, sizes // 10, [8, 8, 32], [0.1] * 3, "float32", True, device="gfx950")


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    assert autotune.lookup("w", path) is None
    best = dict(leaf=32, kernel="stack", queries_per_wave=64, stack_depth=32, ms=1.5, knn_ms=1.0, ball_ms=0.5,
                depth=20)
    autotune.save("w", best, path)
    autotune.save("v", dict(best, leaf=8), path)
    got = autotune.lookup("w", path)
    assert got["leaf"] == 32 and got["kernel"] == "stack" and got["stack_depth"] == 32 and "depth" not in got
    assert autotune.lookup("v", path)["leaf"] == 8
