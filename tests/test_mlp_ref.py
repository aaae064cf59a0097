"""Pins the float64 oracle of the dispatcher network: its analytic gradient
must match central finite differences, and its features must match values
worked out by hand."""
import numpy as np

from oracle import mlp_ref as M


def _rand(seed, B=7):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, M.NP)
    X = rng.uniform(-1, 1, (B, M.NF))
    T = rng.normal(0, 1, (B, M.NC))
    Mask = (rng.uniform(0, 1, (B, M.NC)) > 0.3).astype(np.float64)
    Mask[0, 0] = 1.0
    return p, X, T, Mask


def test_gradient_matches_finite_differences():
    p, X, T, Mask = _rand(1)
    loss, g = M.grad(p, X, T, Mask)
    assert abs(loss - M.loss_only(p, X, T, Mask)) < 1e-12
    rng = np.random.default_rng(2)
    idx = rng.choice(M.NP, 200, replace=False)
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_only(pp, X, T, Mask) - M.loss_only(pm, X, T, Mask)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_forward_is_relu_mlp():
    p = np.zeros(M.NP)
    layers = M.unpack(p)
    assert [W.shape for W, _ in layers] == [(32, 12), (32, 32), (32, 32), (4, 32)]
    # only the last bias set: the output is that bias for any input
    p[-4:] = [1.0, -2.0, 3.0, 0.5]
    y, _ = M.forward(p, np.ones((3, M.NF)))
    assert np.allclose(y, [[1.0, -2.0, 3.0, 0.5]] * 3)


def test_features_by_hand():
    n = 1000
    # four removals, all in the red zone [996, 1000), consecutive and ascending
    f = M.features(np.array([996, 997, 998, 999]), n, 8)
    assert f[5] == 1.0 and f[8] == 1.0 and f[9] == 1.0 and f[10] == 1.0
    assert abs(f[2] - 0.004) < 1e-12 and f[4] == 0.5 and f[11] == 1.0
    assert abs(f[6] - 997.5 / 1000) < 1e-12
    # descending, far apart, none in the red zone
    f = M.features(np.array([900, 500, 100]), n, 4)
    assert f[5] == 0.0 and f[8] == 0.0 and f[9] == 0.0 and f[10] == 0.0
    # empty removal list
    f = M.features(np.array([], np.int64), n, 16)
    assert f[1] == 0.0 and f[3] == -1.0 and f[5] == 0.0
