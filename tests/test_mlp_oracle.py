"""Pin the MLP oracle (oracle/mlp_oracle.py) by mathematics: its backward
pass must equal central finite differences of its own loss when the bf16
roundings are switched off, and its features must follow their definition
on a hand-made grid."""
import numpy as np

from oracle import mlp_oracle as M


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1).astype(np.float64)
    p[M.OFF["b1"][0]:M.OFF["b1"][1]] = rng.standard_normal(64) * 0.1
    p[M.OFF["b3"][0]] = 0.3
    X = rng.standard_normal((40, M.NIN))
    t = rng.standard_normal(40)
    _, g = M.loss_and_grad(p, X, t, rounding=False)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [M.OFF["b3"][0], M.OFF["W1"][0] + 5, M.OFF["w3"][0] + 7]])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_and_grad(pp, X, t, rounding=False)[0] - M.loss_and_grad(pm, X, t, rounding=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_planted_backward_error_is_caught():
    """A transposed W2 in the backward pass must fail the same check."""
    rng = np.random.default_rng(1)
    p = M.init_params(2).astype(np.float64)
    X = rng.standard_normal((30, M.NIN))
    t = rng.standard_normal(30)
    y, c = M.forward(p, X, rounding=False)
    dy = 2 * (y - t) / len(t)
    w = c["w"]
    dz2 = np.outer(dy, w["w3"]) * M.gelu_grad(c["z2"])
    wrong = (dz2 @ w["W2"]) * M.gelu_grad(c["z1"])  # should be W2.T
    right = (dz2 @ w["W2"].T) * M.gelu_grad(c["z1"])
    _, g = M.loss_and_grad(p, X, t, rounding=False)
    gb1 = g[M.OFF["b1"][0]:M.OFF["b1"][1]]
    assert np.allclose(gb1, right.sum(axis=0))
    assert not np.allclose(gb1, wrong.sum(axis=0))


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-3], dtype=np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert r[3] == -2.5


def test_features_definition():
    G, L = 4, 8.0
    grid = np.arange(G ** 3, dtype=np.int32)
    q = np.array([[-4.0, -4.0, -4.0], [1.5, -0.5, 3.9]])
    f = M.features(q, L, grid, G)
    assert np.allclose(f[0, :3], [0, 0, 0]) and np.allclose(f[1, :3], [(1.5 + 4) / 8, 3.5 / 8, 7.9 / 8])
    # query 0 sits in cell (0,0,0); its neighbour (dz,dy,dx)=(-1,-1,-1) wraps to (3,3,3)
    assert np.isclose(f[0, 3], np.log1p(grid[(3 * G + 3) * G + 3]))
    assert np.isclose(f[0, 3 + 13], np.log1p(grid[0]))  # centre cell
    assert np.all(f[:, 30:] == 0)
