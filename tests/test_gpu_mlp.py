"""The HIP feature, forward, gradient and Adam kernels against the float64
oracle (oracle/mlp_ref.py); training reduces the loss; the dispatcher falls
back to the k < n/2 rule without a model."""
import numpy as np
import pytest
import torch

from oracle import mlp_ref as MR
from redzone_amd import dispatch as D
from redzone_amd import workloads as W
from tests.parity import announce

pytestmark = pytest.mark.gpu


def test_features_match_oracle(ctx):
    rng = np.random.default_rng(3)
    lists, host = [], []
    for i in range(40):
        n = int(2 ** rng.uniform(4, 20))
        k = int(rng.integers(0, n + 1))
        esz = int(rng.choice([4, 8, 16]))
        pattern = ["uniform", "clustered", "tail_heavy"][i % 3]
        R = W.removals(pattern, n, k, seed=i)
        A = W.payload(esz, n, seed=i).cuda()
        lists.append((A, R.cuda()))
        host.append((R.numpy(), n, esz))
    announce("features")
    disp = D.Dispatcher(ctx)
    F = disp.features(lists).cpu().numpy()
    for i, (R, n, esz) in enumerate(host):
        want = MR.features(R, n, esz)
        assert np.allclose(F[i], want, atol=2e-6, rtol=1e-5), (i, F[i], want)


def test_forward_and_grad_match_oracle(ctx):
    rng = np.random.default_rng(4)
    p = (rng.normal(0, 0.3, MR.NP)).astype(np.float32)
    X = rng.uniform(-1, 1, (300, MR.NF)).astype(np.float32)
    T = rng.normal(0, 1, (300, MR.NC)).astype(np.float32)
    Mk = (rng.uniform(0, 1, (300, MR.NC)) > 0.2).astype(np.float32)
    disp = D.Dispatcher(ctx, params=p)
    Xd, Td, Md = (torch.from_numpy(a).cuda() for a in (X, T, Mk))
    announce("forward")
    Y, C = disp.predict(Xd)
    y_ref, _ = MR.forward(p, X)
    assert np.allclose(Y.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-4)
    assert np.array_equal(C.cpu().numpy(), np.argmin(Y.cpu().numpy(), axis=1))
    announce("grad")
    loss, g = disp.grad(Xd, Td, Md)
    loss_ref, g_ref = MR.grad(p, X, T, Mk)
    assert abs(float(loss.item()) - loss_ref) <= 1e-4 * abs(loss_ref) + 1e-6
    g = g.cpu().numpy()
    scale = np.abs(g_ref).max()
    assert np.abs(g - g_ref).max() <= 1e-4 * scale, np.abs(g - g_ref).max() / scale
    # deterministic: the same gradient bits again
    _, g2 = disp.grad(Xd, Td, Md)
    assert np.array_equal(g, g2.cpu().numpy())


def test_training_reduces_loss(ctx):
    """Fit a smooth teacher function of the features with the HIP Adam steps."""
    rng = np.random.default_rng(5)
    X = rng.uniform(0, 1, (2048, MR.NF)).astype(np.float32)
    X[:, 11] = 1.0
    Wt = rng.normal(0, 1, (MR.NF, MR.NC))
    T = (np.sin(X @ Wt) + X[:, :4] @ rng.normal(0, 1, (4, MR.NC))).astype(np.float32)
    Mk = np.ones_like(T)
    disp = D.Dispatcher(ctx, seed=1)
    Xd, Td, Md = (torch.from_numpy(a).cuda() for a in (X, T, Mk))
    announce("train")
    loss0 = float(disp.grad(Xd, Td, Md)[0].item())
    disp.train(Xd, Td, Md, epochs=60, lr=3e-3, batch=256, log_every=20)
    loss1 = float(disp.grad(Xd, Td, Md)[0].item())
    assert loss1 < loss0 / 10, (loss0, loss1)
    # the device parameters still match the oracle's forward pass
    y_ref, _ = MR.forward(disp.params.cpu().numpy(), X[:64])
    Y, _ = disp.predict(Xd[:64].contiguous())
    assert np.allclose(Y.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-4)


def test_untrained_dispatcher_uses_rule(ctx):
    disp = D.Dispatcher(ctx)
    ch = disp.choose(None, n=[100, 100, 100], k=[10, 50, 90])
    assert list(ch) == [D.RULE_REDZONE, D.RULE_PARTITION, D.RULE_PARTITION]
