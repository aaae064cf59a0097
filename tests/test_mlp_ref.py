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


def test_features_scales_and_spread_by_hand():
    """f[0] = log2 n / 32, f[1] = log2 k / 32, f[3] = log2(k/n) / 20 and f[7],
    the population standard deviation of position / n, on lists worked out by
    hand; and the strided 256-entry sample of a longer list."""
    n = 1000
    f = M.features(np.array([996, 997, 998, 999]), n, 8)
    assert abs(f[0] - 9.965784284662087 / 32) < 1e-12  # log2(1000)
    assert abs(f[1] - 2 / 32) < 1e-12
    assert abs(f[3] - (-7.965784284662087) / 20) < 1e-12  # log2(4/1000)
    # deviations from 0.9975: -0.0015, -0.0005, 0.0005, 0.0015 -> variance 1.25e-6
    assert abs(f[7] - 1.25e-6 ** 0.5) < 1e-12
    f = M.features(np.array([900, 500, 100]), n, 4)
    assert abs(f[1] - 1.584962500721156 / 32) < 1e-12  # log2(3)
    assert abs(f[3] - (-8.380821783940931) / 20) < 1e-12  # log2(3/1000)
    # positions 0.9, 0.5, 0.1: mean 0.5, variance (0.16 + 0 + 0.16) / 3
    assert abs(f[6] - 0.5) < 1e-12 and abs(f[7] - (0.32 / 3) ** 0.5) < 1e-12
    # k = 512 > 256: the sample is every other entry, R[0], R[2], ..., R[510]
    f = M.features(np.arange(512), 1024, 16)
    assert abs(f[0] - 10 / 32) < 1e-12 and abs(f[1] - 9 / 32) < 1e-12 and abs(f[3] - (-1 / 20)) < 1e-12
    assert abs(f[6] - 255 / 1024) < 1e-12
    assert abs(f[7] - 2 * ((256 ** 2 - 1) / 12) ** 0.5 / 1024) < 1e-12  # 2 x std of 0..255
    assert f[5] == 0.0 and f[8] == 1.0 and f[9] == 1.0 and f[10] == 1.0
