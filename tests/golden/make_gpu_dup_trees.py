"""Regression fixture: GPU trees of inputs with many equal coordinates.

With equal coordinates, a slide (every point on one side of the midpoint)
moves the first extreme point in the build's current point order, and the
GPU's order differs from the reference's Lomuto order (the parity tests
compare trees node for node only on duplicate-free inputs).  These trees
pin the GPU build's own choice, so that a change to how the build tracks
extremes and positions keeps every tree node for node.  Made by the GPU
build (level-synchronous partition with a separate extremes pass per
level) before the extremes moved into the partition pass:

    python tests/golden/make_gpu_dup_trees.py   # on a machine with the GPU

writes tests/golden/gpu_dup_trees.npz.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("axis", "split", "lo", "hi", "child", "parent", "start", "count", "idx")


def cases():
    """name -> (points, leaf, boxsize or None)."""
    out = {}
    rng = np.random.default_rng(41)
    # periodic float32 box of side 100 on a 0.5-spaced lattice of 20^3 sites
    g = rng.integers(-20, 20, (8000, 3)) * 2.5
    out["lattice_p3_f32_leaf16"] = (g.astype(np.float32), 16, np.full(3, 100.0))
    # non-periodic float64 integer grid 0..9 with repeated sites
    out["grid_u3_f64_leaf4"] = (rng.integers(0, 10, (6000, 3)).astype(np.float64), 4, None)
    # 2-D float32, leaf 1: half the points are copies of 40 sites
    base = rng.random((1500, 2)).astype(np.float32)
    sites = rng.random((40, 2)).astype(np.float32)
    out["copies_u2_f32_leaf1"] = (np.concatenate([base, sites[rng.integers(0, 40, 1500)]]), 1, None)
    # periodic float64 with one axis constant (ties on every split of it)
    p = (rng.random((5000, 3)) - 0.5) * 50.0
    p[:, 1] = 3.0
    out["flat_p3_f64_leaf8"] = (p, 8, np.full(3, 50.0))
    return out


def main():
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from cosmokdtree_amd import kdtree
    arrays = {}
    for name, (pts, leaf, L) in cases().items():
        t = kdtree._build(pts, leaf, L)
        d = t.dump()
        t.free()
        for f in FIELDS:
            arrays[f"{name}/{f}"] = d[f]
        print(name, "nodes", len(d["axis"]), flush=True)
    np.savez_compressed(os.path.join(HERE, "gpu_dup_trees.npz"), **arrays)


if __name__ == "__main__":
    main()
