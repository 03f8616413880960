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


def test_tree_features_definition():
    """(1/3) ln(count / volume) of the tight box of the >=16 / >=64 ancestor:
    with fewer than 16 points the root is that ancestor (checked by brute
    force); scaling points and box by 2 (exact in floating point, same tree)
    shifts every feature by -ln 2."""
    from oracle.oracle import OracleTree
    rng = np.random.default_rng(3)
    pts = (rng.random((12, 3)) - 0.5) * 10
    orc = OracleTree(pts, 4, np.full(3, 10.0))
    try:
        tf = M.tree_features(orc.dump(), pts[:5], 10.0)
    finally:
        orc.close()
    ext = pts.max(0) - pts.min(0)
    assert np.allclose(tf, np.log(12 / np.prod(ext)) / 3, rtol=0, atol=1e-6)
    pts = (rng.random((3000, 3)) - 0.5) * 10
    q = pts[:300] + rng.normal(0, 0.05, (300, 3))
    q = np.clip(q, -5, 5)
    o1, o2 = OracleTree(pts, 16, np.full(3, 10.0)), OracleTree(pts * 2, 16, np.full(3, 20.0))
    try:
        f1 = M.tree_features(o1.dump(), q, 10.0)
        f2 = M.tree_features(o2.dump(), q * 2, 20.0)
    finally:
        o1.close()
        o2.close()
    assert np.allclose(f2 - f1, -np.log(2), atol=2e-6)
    assert np.all(f1[:, 1] <= f1[:, 0] + 1.5)  # larger ancestors are not much denser


def test_gelu_sigmoid_form():
    """z * sigmoid(1.702 z): exact at 0 and in the tails, within 0.021 of
    the erf form of GELU everywhere, and its derivative by differences."""
    from scipy.special import erf
    z = np.linspace(-6, 6, 2001)
    exact = z * 0.5 * (1 + erf(z / np.sqrt(2)))
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu(np.array([1.0]))[0] - 1 / (1 + np.exp(-1.702))) < 1e-15
    assert np.max(np.abs(M.gelu(z) - exact)) < 0.021
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.allclose(M.gelu_grad(z), fd, atol=1e-7)


def test_adam_three_steps_closed_form():
    """Three Adam steps against the moments written out by hand (Kingma & Ba
    2015, bias-corrected): a constant gradient g gives m_hat = g and
    v_hat = g^2 at every step, so each step moves by lr * g / (|g| + eps);
    the gradients g, -g, g give m = 0.1 g, -0.01 g, 0.091 g with m_hat =
    m / (1 - 0.9^t) and v_hat = g^2."""
    lr, eps = 1e-2, 1e-8
    g = np.array([0.5, -2.0, 3e-3], np.float32)
    p0 = np.array([1.0, -1.0, 0.25], np.float32)
    p, m, v = p0.copy(), np.zeros(3, np.float32), np.zeros(3, np.float32)
    for t in (1, 2, 3):
        p, m, v = M.adam(p, g, m, v, t, lr)
    assert np.allclose(p, p0 - 3 * lr * g / (np.abs(g) + eps), rtol=0, atol=2e-6)
    assert np.allclose(m, (1 - 0.9 ** 3) * g, rtol=1e-5)
    assert np.allclose(v, (1 - 0.999 ** 3) * g * g, rtol=1e-4)
    p, m, v = p0.copy(), np.zeros(3, np.float32), np.zeros(3, np.float32)
    want = p0.astype(np.float64)
    for t, (gt, mt) in enumerate(((g, 0.1), (-g, -0.01), (g, 0.091)), start=1):
        p, m, v = M.adam(p, gt, m, v, t, lr)
        mhat = mt * g.astype(np.float64) / (1 - 0.9 ** t)
        want = want - lr * mhat / (np.abs(g.astype(np.float64)) + eps)
        assert np.allclose(m, mt * g, rtol=1e-5, atol=1e-9), t
    assert np.allclose(p, want, rtol=0, atol=2e-6)
    # a planted mistake (no bias correction) is caught
    pb = p0.astype(np.float64) - lr * 0.1 * g / (np.sqrt(0.001) * np.abs(g) + eps)
    assert not np.allclose(pb, p0 - lr * g / (np.abs(g) + eps), atol=2e-6)
