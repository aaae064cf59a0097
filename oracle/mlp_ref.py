"""CPU oracle (float64 numpy) of the dispatcher's features and MLP.

TEST INFRASTRUCTURE ONLY: the tests compare the HIP feature, forward and
gradient kernels (rz_mlp.hip) with these functions; nothing in the product
imports this file.  Layer sizes and parameter layout follow include/redzone.h:
F=12 -> 32 -> 32 -> 32 -> 4, ReLU, row-major W[out][in] then b[out] per layer.
"""
import numpy as np

NF, NH, NC = 12, 32, 4
NP = NH * NF + NH + 2 * (NH * NH + NH) + NC * NH + NC


def unpack(params):
    p = np.asarray(params, np.float64)
    out, o = [], 0
    for nin, nout in [(NF, NH), (NH, NH), (NH, NH), (NH, NC)]:
        W = p[o: o + nin * nout].reshape(nout, nin)
        o += nin * nout
        b = p[o: o + nout]
        o += nout
        out.append((W, b))
    assert o == NP
    return out


def forward(params, X):
    """Returns (y, activations); activations[0] = X, activations[i] = (pre, post) of layer i."""
    layers = unpack(params)
    h = np.asarray(X, np.float64)
    acts = [h]
    for i, (W, b) in enumerate(layers):
        a = h @ W.T + b
        h = np.maximum(a, 0.0) if i < 3 else a
        acts.append((a, h))
    return h, acts


def grad(params, X, T, M):
    """Loss sum m (y - t)^2 / sum m and its analytic gradient (flat, library layout)."""
    layers = unpack(params)
    T = np.asarray(T, np.float64)
    M = np.asarray(M, np.float64)
    msum = M.sum()
    y, acts = forward(params, X)
    loss = float((M * (y - T) ** 2).sum() / msum)
    d = 2.0 * M * (y - T) / msum
    grads = []
    for i in range(3, -1, -1):
        W, _ = layers[i]
        h_in = acts[0] if i == 0 else acts[i][1]
        grads.append((d.T @ h_in, d.sum(0)))
        if i > 0:
            d = (d @ W) * (acts[i][0] > 0)
    grads.reverse()
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def loss_only(params, X, T, M):
    y, _ = forward(params, X)
    M = np.asarray(M, np.float64)
    return float((M * (y - np.asarray(T, np.float64)) ** 2).sum() / M.sum())


def features(R, n, elem_bytes):
    """The 12 list features of rz_features for removal list R (numpy int array) of a list of n items."""
    R = np.asarray(R, np.int64)
    k = R.shape[0]
    sz = n - k
    m = min(k, 256)
    f = np.zeros(NF)
    f[0] = np.log2(max(n, 1)) / 32
    f[1] = np.log2(k) / 32 if k > 0 else 0.0
    f[2] = k / max(n, 1)
    f[3] = max(np.log2(k / n) / 20, -1.0) if k > 0 else -1.0
    f[4] = elem_bytes / 16
    if m > 0:
        idx = (np.arange(m, dtype=np.int64) * k) // m
        s = R[idx]
        x = s / float(max(n, 1))
        f[5] = (s >= sz).mean()
        f[6] = x.mean()
        f[7] = np.sqrt(max((x * x).mean() - x.mean() ** 2, 0.0))
        if m > 1:
            f[10] = (x[1:] > x[:-1]).mean()
    mp = 0 if k < 2 else min(k - 1, 128)
    if mp > 0:
        i = (np.arange(mp, dtype=np.int64) * (k - 1)) // mp
        a, b = R[i], R[i + 1]
        f[8] = (b == a + 1).mean()
        f[9] = (np.abs(b - a) < 64).mean()
    f[11] = 1.0
    return f
