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


def _mixed_lists(seed, count):
    rng = np.random.default_rng(seed)
    lists = []
    for i in range(count):
        n = int(2 ** rng.uniform(4, 18))
        k = [0, n][i % 2] if i % 7 == 0 else int(rng.integers(1, n))
        esz = int(rng.choice([4, 8, 16]))
        R = W.removals(["uniform", "clustered", "tail_heavy"][i % 3], n, k, seed=seed + i)
        lists.append((W.payload(esz, n, seed=i).cuda(), R.cuda()))
    return lists


def test_batched_choice_matches_forward(ctx):
    """rz_dispatch_choose / rz_dispatch_collect (one call for many lists)
    against the features and forward pass called separately: the same argmin
    for every list with 0 < k < n, Redzone for k = 0 and k = n; several calls
    in a row, either side of the 48 lists that one kernel takes as its
    argument (above, the descriptors are uploaded), with this dispatcher's
    parameters and with the ones installed in the context (the committed
    model in both)."""
    disp = D.Dispatcher.load(ctx)
    assert disp.trained
    for seed, count in ((90, 70), (91, 23), (93, 48), (94, 49)):  # 48: the most lists passed as a kernel argument
        lists = _mixed_lists(seed, count)
        descs = ctx.descs(lists)
        announce("batched choice", seed, count)
        want = disp.predict(disp.features(descs))[1].cpu().numpy()
        n, k = descs.rows["n"], descs.rows["k"]
        want = np.where((k == 0) | (k == n), D.RULE_REDZONE, want)
        disp.choose_async(descs)
        got = disp.collect()
        assert np.array_equal(got, want), (got, want)
        # null parameters: the model installed in the context
        out = np.empty(count, np.int32)
        ctx._call("rz_dispatch_choose", ctx.handle, None, descs.as_ctypes(), count, D._stream_ptr(None, ctx.device))
        ctx._call("rz_dispatch_collect", ctx.handle, out.ctypes.data_as(D.ctypes.POINTER(D.ctypes.c_int32)), count)
        assert np.array_equal(out, want)


def test_batched_choice_untrained_uses_rule(ctx):
    descs = ctx.descs(_mixed_lists(92, 30))
    disp = D.Dispatcher(ctx)
    disp.choose_async(descs)
    n, k = descs.rows["n"], descs.rows["k"]
    want = np.where((k == 0) | (k == n), D.RULE_REDZONE, np.where(2 * k < n, D.RULE_REDZONE, D.RULE_PARTITION))
    assert np.array_equal(disp.collect(), want)
