"""Learned dispatch between Redzone launch forms and the O(n) partition.

A small FP32 MLP (rz_mlp.hip: 12 features -> 32 -> 32 -> 32 -> 4) predicts
log2 of the time in microseconds of each method for a list; the dispatcher
runs the method with the smallest prediction.  Without a model it falls back
to the fixed rule of the reference: Redzone when k < n/2, else the O(n)
partition (README.md:6; paper p.14).

Features, forward pass, gradients and Adam steps all run in the library's HIP
kernels; this module only drives them (collects timings, batches the
training, stores the parameters).

Methods (outputs of the network):
  0  redzone         rz_remove with the tuned launch shape for (element size, log2 k)
  1  redzone_single  one ordered launch (roles by workgroup id)
  2  redzone_dense   the region-bucketed form (variant 5), which pays off when
                     k/n is large enough that holes share DRAM bursts
  3  partition       rz_partition_on (mark + stable O(n) compaction)
"""
import ctypes
import math
import os

import numpy as np

from . import _native
from ._native import ListDesc, TileConfig

NF = 12
NH = 32
NC = 4
NP = NH * NF + NH + 2 * (NH * NH + NH) + NC * NH + NC
assert NP == 2660

METHODS = ["redzone", "redzone_single", "redzone_dense", "partition"]
RULE_REDZONE, RULE_PARTITION = 0, 3


def method_config(method, elem_bytes):
    """Launch shape of a Redzone method (None: the tuned table), as the library
    defines it (rz_dispatch_method_config)."""
    if method not in (1, 2):
        return None
    out = TileConfig()
    st = _native.load().rz_dispatch_method_config(method, elem_bytes, ctypes.byref(out))
    if st:
        raise ValueError(f"no launch shape for method {method}, {elem_bytes}-byte items")
    return out

MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dispatch_model_gfx950.npz")


def rule_choice(n, k):
    """The reference's fixed rule: Redzone is faster than O(n) when k < n/2."""
    return RULE_REDZONE if 2 * k < n else RULE_PARTITION


def run_method(ctx, method, data, removals, stream=None):
    """Runs method `method` on one list (in place); returns n - k."""
    if method == 3:
        return ctx.partition(data, removals, stream=stream)
    from . import elem_bytes_of

    return ctx.remove(data, removals, config=method_config(method, elem_bytes_of(data)), stream=stream)


def init_params(seed=0):
    """He-uniform initial parameters in the library's layout (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for nin, nout in [(NF, NH), (NH, NH), (NH, NH), (NH, NC)]:
        lim = math.sqrt(6.0 / nin)
        parts.append(rng.uniform(-lim, lim, size=(nout, nin)).astype(np.float32).ravel())
        parts.append(np.zeros(nout, np.float32))
    p = np.concatenate(parts)
    assert p.size == NP
    return p


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Dispatcher:
    """MLP dispatcher whose parameters live on the device."""

    def __init__(self, ctx, params=None, seed=0):
        import torch

        self.ctx = ctx
        self.lib = _native.load()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.params = torch.from_numpy(p.copy()).cuda(ctx.device)
        self.trained = params is not None
        self.train_log = []
        self._pending = None  # the choice issued by choose_async

    # ---- features -------------------------------------------------------
    def features(self, lists, stream=None):
        """lists: [(data, removals), ...] device tensors, or a ListDescs -> float32 device tensor [B, NF]."""
        import torch

        from . import ListDescs

        if not isinstance(lists, ListDescs):
            lists = self.ctx.descs(lists)
        cnt = len(lists)
        d_descs = torch.from_numpy(lists.rows.view(np.uint8)).cuda(self.ctx.device)
        out = torch.empty((cnt, NF), dtype=torch.float32, device=f"cuda:{self.ctx.device}")
        self.ctx._call("rz_features", self.ctx.handle, _ptr(d_descs), cnt, _ptr(out),
                       _stream_ptr(stream, self.ctx.device))
        if stream is not None:
            d_descs.record_stream(stream)  # keep the upload alive until the kernel has read it
        return out

    # ---- inference ------------------------------------------------------
    def predict(self, X, stream=None):
        import torch

        B = X.shape[0]
        Y = torch.empty((B, NC), dtype=torch.float32, device=X.device)
        C = torch.empty(B, dtype=torch.int32, device=X.device)
        self.ctx._call("rz_mlp_forward", self.ctx.handle, _ptr(self.params), _ptr(X.contiguous()), B, _ptr(Y), _ptr(C),
                       _stream_ptr(stream, self.ctx.device))
        return Y, C

    def choose(self, X, n=None, k=None):
        """Method index per row; the k < n/2 rule when the network is untrained."""
        if not self.trained:
            return np.array([rule_choice(a, b) for a, b in zip(n, k)], np.int32)
        _, C = self.predict(X)
        return C.cpu().numpy()

    def choose_async(self, lists, stream=None):
        """Issues the method choice for a ListDescs (rz_dispatch_choose, with
        this dispatcher's parameters) on the stream without waiting; collect()
        returns it.  An untrained dispatcher applies the k < n/2 rule."""
        if not self.trained:
            n, k = lists.rows["n"], lists.rows["k"]
            rule = np.array([rule_choice(a, b) for a, b in zip(n, k)], np.int32)
            self._pending = np.where((k == 0) | (k == n), RULE_REDZONE, rule).astype(np.int32)
            return
        self._pending = len(lists)
        self.ctx._call("rz_dispatch_choose", self.ctx.handle, _ptr(self.params), lists.as_ctypes(), len(lists),
                       _stream_ptr(stream, self.ctx.device))

    def collect(self):
        """The choice issued by choose_async (int32 RZ_METHOD_* per list)."""
        if isinstance(self._pending, np.ndarray):
            out, self._pending = self._pending, None
            return out
        out = np.empty(self._pending, np.int32)
        self.ctx._call("rz_dispatch_collect", self.ctx.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                       len(out))
        self._pending = None
        return out

    # ---- training -------------------------------------------------------
    def grad(self, X, T, M):
        """Loss and gradient of the masked MSE on log2 times (device tensors)."""
        import torch

        B = X.shape[0]
        grad = torch.empty(NP, dtype=torch.float32, device=X.device)
        loss = torch.empty(1, dtype=torch.float32, device=X.device)
        work = torch.empty(B * (NP + 1), dtype=torch.float32, device=X.device)
        msum = float(M.sum().item())
        self.ctx._call("rz_mlp_grad", self.ctx.handle, _ptr(self.params), _ptr(X.contiguous()), _ptr(T.contiguous()),
                       _ptr(M.contiguous()), B, ctypes.c_float(msum), _ptr(grad), _ptr(loss), _ptr(work),
                       _stream_ptr(None, self.ctx.device))
        return loss, grad

    def train(self, X, T, M, epochs=400, lr=3e-3, batch=256, seed=0, log_every=50):
        """Adam on mini-batches; X [N, NF], T/M [N, NC] device float32 tensors."""
        import torch

        N = X.shape[0]
        m = torch.zeros(NP, dtype=torch.float32, device=X.device)
        v = torch.zeros(NP, dtype=torch.float32, device=X.device)
        rng = np.random.default_rng(seed)
        step = 0
        for ep in range(epochs):
            order = torch.from_numpy(rng.permutation(N)).to(X.device)
            for s in range(0, N, batch):
                idx = order[s: s + batch]
                Xb, Tb, Mb = X[idx].contiguous(), T[idx].contiguous(), M[idx].contiguous()
                if float(Mb.sum().item()) <= 0:
                    continue
                _, g = self.grad(Xb, Tb, Mb)
                step += 1
                self.ctx._call("rz_mlp_adam", self.ctx.handle, _ptr(self.params), _ptr(g), _ptr(m), _ptr(v), step,
                               ctypes.c_float(lr), _stream_ptr(None, self.ctx.device))
            if log_every and (ep % log_every == 0 or ep == epochs - 1):
                loss, _ = self.grad(X, T, M)
                self.train_log.append((ep, float(loss.item())))
        self.trained = True
        return self.train_log

    # ---- persistence ----------------------------------------------------
    def save(self, path=MODEL_PATH, meta=None):
        np.savez(path, params=self.params.cpu().numpy(), meta=np.array(repr(meta or {})))

    @classmethod
    def load(cls, ctx, path=MODEL_PATH):
        with np.load(path, allow_pickle=False) as z:
            return cls(ctx, params=z["params"])


def _stream_ptr(stream, device):
    import torch

    if stream is None:
        stream = torch.cuda.current_stream(device)
    return ctypes.c_void_p(stream.cuda_stream)
