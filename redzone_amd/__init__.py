"""redzone_amd -- Redzone O(k) in-place stream compaction on MI355X (gfx950).

Host-side mirror of the reference's entry point shape
`remove(data, n, removeList, k)` (RemoveFromList_GPUWide(T* List, int Length,
int* Removals, int RemoveCount, ...), redzone.hxx:626-941) over the C ABI of
librz_amd.so (include/redzone.h).  Tensors are torch device tensors used as
plain device memory; every computation happens in the library's HIP kernels.

    ctx = Context(device=0)
    new_len = ctx.remove(data, removals)          # O(k) Redzone, in place
    new_len = ctx.partition(data, removals)       # O(n) stable partition
"""
import ctypes

from . import _native
from ._native import (RZ_COOPERATIVE, RZ_PRESERVE_REMOVALS, RZ_PROFILE, RZ_STATS, RZ_VALIDATE, RedzoneError,
                      Stats, TileConfig)

__all__ = ["Context", "remove", "partition", "TileConfig", "Stats", "RedzoneError", "RZ_PRESERVE_REMOVALS",
           "RZ_VALIDATE", "RZ_COOPERATIVE", "RZ_STATS", "RZ_PROFILE", "elem_bytes_of", "default_context"]


def elem_bytes_of(data):
    """Bytes per item of a tensor of shape (n,) or (n, w): the whole row is one item."""
    per_row = 1
    for d in data.shape[1:]:
        per_row *= d
    return per_row * data.element_size()


def _stream_handle(stream, device):
    import torch

    if stream is None:
        stream = torch.cuda.current_stream(device)
    return ctypes.c_void_p(stream.cuda_stream)


def _check_tensor(t, name, device):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.device.index != device:
        raise ValueError(f"{name} is on {t.device}, the context is on cuda:{device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class Context:
    """Owns an rz_ctx (device scratch + tuning table) on one device."""

    def __init__(self, device=0):
        self._lib = _native.load()
        self.device = int(device)
        h = ctypes.c_void_p()
        st = self._lib.rz_ctx_create(ctypes.byref(h), self.device)
        if st != 0:
            raise RedzoneError(st, "rz_ctx_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rz_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def device_info(self):
        cu = ctypes.c_int32()
        arch = ctypes.create_string_buffer(64)
        self._call("rz_device_info", self._h, ctypes.byref(cu), arch, 64)
        return {"cu_count": cu.value, "arch": arch.value.decode()}

    def _call(self, name, *args):
        st = getattr(self._lib, name)(*args)
        if st != 0:
            raise RedzoneError(st, name)
        return st

    def _args(self, data, removals, n, k):
        _check_tensor(data, "data", self.device)
        _check_tensor(removals, "removals", self.device)
        import torch

        if removals.dtype != torch.int32:
            raise ValueError("removals must be int32")
        n = data.shape[0] if n is None else int(n)
        k = removals.shape[0] if k is None else int(k)
        if n > data.shape[0] or k > removals.shape[0]:
            raise ValueError("n / k exceed the tensors")
        return n, k, elem_bytes_of(data)

    def remove(self, data, removals, n=None, k=None, flags=0, config=None, stream=None):
        """Redzone O(k) removal in place; returns the new length n - k.

        data[:n-k] holds the survivors afterwards; data[n-k:n] and removals
        (unless RZ_PRESERVE_REMOVALS) are scratch."""
        n, k, esz = self._args(data, removals, n, k)
        s = _stream_handle(stream, self.device)
        if config is None:
            self._call("rz_remove", self._h, ctypes.c_void_p(data.data_ptr()), n, esz,
                       ctypes.c_void_p(removals.data_ptr()), k, flags, s)
        else:
            self._call("rz_remove_config", self._h, ctypes.c_void_p(data.data_ptr()), n, esz,
                       ctypes.c_void_p(removals.data_ptr()), k, flags, ctypes.byref(config), s)
        return n - k

    def partition(self, data, removals, n=None, k=None, flags=0, stream=None):
        """O(n) stable in-place partition; survivors keep their order.  Returns n - k."""
        n, k, esz = self._args(data, removals, n, k)
        s = _stream_handle(stream, self.device)
        self._call("rz_partition_on", self._h, ctypes.c_void_p(data.data_ptr()), n, esz,
                   ctypes.c_void_p(removals.data_ptr()), k, flags, s)
        return n - k

    def last_profile(self):
        """Kernel times (ms) of the last RZ_PROFILE call: [mark, move, pair, total]."""
        ms = (ctypes.c_float * 4)()
        self._call("rz_last_profile", self._h, ms)
        return list(ms)

    def last_stats(self):
        out = Stats()
        self._call("rz_last_stats", self._h, ctypes.byref(out))
        return out


_default = {}


def default_context(device=0):
    if device not in _default:
        _default[device] = Context(device)
    return _default[device]


def remove(data, n, removeList, k, ctx=None, flags=0, stream=None):
    """remove(data, n, removeList, k): the reference's entry-point shape."""
    ctx = ctx or default_context(data.device.index)
    return ctx.remove(data, removeList, n=n, k=k, flags=flags, stream=stream)


def partition(data, n, removeList, k, ctx=None, flags=0, stream=None):
    ctx = ctx or default_context(data.device.index)
    return ctx.partition(data, removeList, n=n, k=k, flags=flags, stream=stream)
