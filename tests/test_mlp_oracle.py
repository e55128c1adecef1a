"""The CPU restatement of the learned split-cost MLP, pinned by finite
differences (float64 form), and the shape of its FP16 form."""
import numpy as np

from oracle import mlp


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp.NPARAM)
    p[mlp.OFF_W1:mlp.OFF_B1] = rng.uniform(-1, 1, mlp.H * mlp.IN) * np.sqrt(6 / mlp.IN)
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.uniform(-0.2, 0.2, mlp.H)
    p[mlp.OFF_W2:mlp.OFF_B2] = rng.uniform(-1, 1, mlp.H * mlp.H) * np.sqrt(6 / mlp.H)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.uniform(-0.2, 0.2, mlp.H)
    p[mlp.OFF_W3:mlp.OFF_B3] = rng.uniform(-1, 1, mlp.H) * np.sqrt(3 / mlp.H)
    p[mlp.OFF_B3] = -3.0
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rand_params(2)
    X = rng.uniform(-1, 1, (50, mlp.IN))
    T = rng.normal(size=50)
    W = rng.uniform(0.5, 2, 50)
    _, g = mlp.loss_and_grad(p, X, T, W, quantize=False)
    h = 1e-6
    for i in list(range(0, mlp.NPARAM, 7)) + [mlp.OFF_B3]:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (mlp.loss_and_grad(pp, X, T, W, quantize=False)[0] - mlp.loss_and_grad(pm, X, T, W, quantize=False)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), (i, num, g[i])


def test_fp16_form_stays_close_to_float64():
    rng = np.random.default_rng(3)
    p = rand_params(4)
    X = rng.uniform(-1, 1, (200, mlp.IN))
    y64, _ = mlp.forward(p, X, quantize=False)
    y16, _ = mlp.forward(p, X, quantize=True)
    assert np.abs(y64 - y16).max() < 0.05 * (1 + np.abs(y64).max())
    assert not np.array_equal(y64, y16)


def test_gradient_descent_on_a_known_function():
    """Plain gradient steps on the oracle's own gradient fit a smooth target."""
    rng = np.random.default_rng(5)
    p = rand_params(6)
    X = rng.uniform(0, 1, (256, mlp.IN))
    T = -4 + 2 * X[:, 0] - X[:, 6]
    l0, _ = mlp.loss_and_grad(p, X, T)
    for _ in range(300):
        _, g = mlp.loss_and_grad(p, X, T, quantize=False)
        p -= 0.02 * g
    l1, _ = mlp.loss_and_grad(p, X, T)
    assert l1 < 0.1 * l0


def test_split_features_closed_form():
    # child [0.25,0,0]-[0.75,1,0.5] with 256 of the 1024 primitives of the
    # parent [0,0,0]-[1,1,1], in the scene box [0,0,0]-[2,2,2]
    f = mlp.split_features([[0.25, 0, 0]], [[0.75, 1, 0.5]], [256], [[0, 0, 0]], [[1, 1, 1]], [1024],
                           [0, 0, 0], [2, 2, 2])[0]
    # half areas: child 0.5*(1+0.5)+1*0.5 = 1.25, parent 3, scene 12
    assert np.isclose(f[0], np.log2(1.25 / 3 + 2.0 ** -30) / 16, rtol=1e-6)
    assert f[1] == 0.25
    assert np.allclose(f[2:5], [0.0, 0.0, -0.25])   # centre offsets over the parent's extent
    assert np.allclose(f[5:8], [0.5, 1.0, 0.5])     # extent ratios
    assert np.allclose(f[8:11], [0.25, 0.25, 0.25])  # parent centre in the scene box
    assert np.isclose(f[11], np.log2(3 / 12 + 2.0 ** -30) / 16, rtol=1e-6)
    # a child equal to its parent: SAH estimate 1 (feature 0 at the floor's
    # distance from 0), no offset, full extent
    g = mlp.split_features([[0, 0, 0]], [[1, 1, 1]], [8], [[0, 0, 0]], [[1, 1, 1]], [8], [0, 0, 0], [1, 1, 1])[0]
    assert abs(g[0]) < 1e-9 and g[1] == 1 and np.allclose(g[2:5], 0) and np.allclose(g[5:8], 1)
    # a flat parent axis: offset 0 and extent ratio 1 on it
    h = mlp.split_features([[0, 0, 0.5]], [[0.5, 1, 0.5]], [1], [[0, 0, 0.5]], [[1, 1, 0.5]], [2], [0, 0, 0],
                           [1, 1, 1])[0]
    assert h[4] == 0 and h[7] == 1 and np.isclose(h[5], 0.5)
