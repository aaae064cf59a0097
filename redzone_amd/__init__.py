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
import os

from . import _native
from ._native import (RZ_COOPERATIVE, RZ_EDEVICE, RZ_ERR_LOOKBACK_TIMEOUT, RZ_ERR_ORPHAN_MISMATCH, RZ_ERR_SIZES,
                      RZ_ERR_WAIT_TIMEOUT,
                      RZ_METHOD_DENSE, RZ_METHOD_ORDERED, RZ_METHOD_PARTITION, RZ_METHOD_REDZONE, RZ_ANY_PAIRING,
                      RZ_PRESERVE_REMOVALS, RZ_PROFILE, RZ_SORTED, RZ_STATS, RZ_VALIDATE, BatchInfo, ListDesc,
                      RedzoneError, Stats, TileConfig)

__all__ = ["Context", "Batch", "ListDescs", "remove", "partition", "TileConfig", "Stats", "RedzoneError", "RZ_PRESERVE_REMOVALS",
           "RZ_VALIDATE", "RZ_COOPERATIVE", "RZ_STATS", "RZ_PROFILE", "RZ_SORTED", "RZ_ANY_PAIRING", "RZ_METHOD_REDZONE",
           "RZ_METHOD_ORDERED", "RZ_METHOD_DENSE", "RZ_METHOD_PARTITION", "RZ_EDEVICE", "RZ_ERR_ORPHAN_MISMATCH",
           "RZ_ERR_WAIT_TIMEOUT", "RZ_ERR_LOOKBACK_TIMEOUT", "RZ_ERR_SIZES", "elem_bytes_of", "default_context"]


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

    def __init__(self, device=0, load_tuned=True, load_dispatcher=True):
        self._lib = _native.load()
        self.device = int(device)
        h = ctypes.c_void_p()
        st = self._lib.rz_ctx_create(ctypes.byref(h), self.device)
        if st != 0:
            raise RedzoneError(st, "rz_ctx_create")
        self._h = h
        self.tuned_table = None
        self.dispatcher_model = None
        gfx950 = self.device_info()["arch"].startswith("gfx950")
        if load_tuned and os.path.exists(_native.TUNED_TABLE) and gfx950:
            self.load_table(_native.TUNED_TABLE)
            self.tuned_table = _native.TUNED_TABLE
        model = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dispatch_model_gfx950.npz")
        if load_dispatcher and os.path.exists(model) and gfx950:
            import numpy as np

            with np.load(model, allow_pickle=False) as z:
                self.set_dispatch_params(z["params"])
            self.dispatcher_model = model

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

    def remove_sized(self, data, removals, d_n, d_k, n_cap=None, k_cap=None, flags=0, stream=None):
        """remove() with n and k read on the device from the one-element int32
        tensors d_n and d_k (TestGPUWide's sizeA / sizeR, redzone.hxx:457-465).
        n_cap (default data.shape[0]) and k_cap (default removals.shape[0])
        bound them; the first k entries of removals are read.  Returns nothing:
        the new length n - k is on the device.  Sizes that break their bounds
        move nothing and set RZ_ERR_SIZES (see check_errors)."""
        import torch

        n_cap, k_cap, esz = self._args(data, removals, n_cap, k_cap)
        for t, name in ((d_n, "d_n"), (d_k, "d_k")):
            _check_tensor(t, name, self.device)
            if t.dtype != torch.int32 or t.numel() < 1:
                raise ValueError(f"{name} must be an int32 tensor of at least one element")
        self._call("rz_remove_sized", self._h, ctypes.c_void_p(data.data_ptr()), n_cap, esz,
                   ctypes.c_void_p(removals.data_ptr()), k_cap, ctypes.c_void_p(d_n.data_ptr()),
                   ctypes.c_void_p(d_k.data_ptr()), flags, _stream_handle(stream, self.device))

    def add_remove(self, data, n, additions, removals, k=None, flags=0, stream=None):
        """Appends `additions` after data[:n] (data must hold n + g items), then
        removes `removals` (indices into the combined list) as remove() would on
        n + g items.  Returns n + g - k."""
        _check_tensor(data, "data", self.device)
        _check_tensor(additions, "additions", self.device)
        _check_tensor(removals, "removals", self.device)
        import torch

        if removals.dtype != torch.int32:
            raise ValueError("removals must be int32")
        esz = elem_bytes_of(data)
        if elem_bytes_of(additions) != esz:
            raise ValueError("additions must have the list's item size")
        g = additions.shape[0]
        k = removals.shape[0] if k is None else int(k)
        self._call("rz_add_remove", self._h, ctypes.c_void_p(data.data_ptr()), int(n), data.shape[0], esz,
                   ctypes.c_void_p(additions.data_ptr()), g, ctypes.c_void_p(removals.data_ptr()), k, flags,
                   _stream_handle(stream, self.device))
        return int(n) + g - k

    def gather_preserve(self, data, keep, n=None, p=None, flags=0, stream=None):
        """In place: data[i] = data[keep[i]] for i < p (keep strictly ascending,
        int32).  Returns p."""
        n, p, esz = self._args(data, keep, n, p)
        self._call("rz_gather_preserve", self._h, ctypes.c_void_p(data.data_ptr()), n, esz,
                   ctypes.c_void_p(keep.data_ptr()), p, flags, _stream_handle(stream, self.device))
        return p

    def mark_removals(self, removals, n, tail_bits, k=None, stream=None):
        """The mark step alone into tail_bits (int32 device tensor of ceil(k/32)
        words, zero beforehand): bit r-(n-k) for every removal r >= n-k."""
        _check_tensor(removals, "removals", self.device)
        _check_tensor(tail_bits, "tail_bits", self.device)
        k = removals.shape[0] if k is None else int(k)
        if tail_bits.numel() * tail_bits.element_size() < 4 * ((k + 31) // 32):
            raise ValueError("tail_bits holds fewer than ceil(k/32) words")
        self._call("rz_mark_removals", self._h, ctypes.c_void_p(removals.data_ptr()), int(n), k,
                   ctypes.c_void_p(tail_bits.data_ptr()), _stream_handle(stream, self.device))

    def remove_premarked(self, data, removals, tail_bits, n=None, k=None, flags=0, config=None, stream=None):
        """remove() with the tail marks already in tail_bits (see mark_removals and
        include/redzone_device.h); leaves tail_bits zero.  Returns n - k."""
        n, k, esz = self._args(data, removals, n, k)
        _check_tensor(tail_bits, "tail_bits", self.device)
        if tail_bits.numel() * tail_bits.element_size() < 4 * ((k + 31) // 32):
            raise ValueError("tail_bits holds fewer than ceil(k/32) words")
        cfg = ctypes.byref(config) if config is not None else None
        self._call("rz_remove_premarked", self._h, ctypes.c_void_p(data.data_ptr()), n, esz,
                   ctypes.c_void_p(removals.data_ptr()), k, ctypes.c_void_p(tail_bits.data_ptr()), flags, cfg,
                   _stream_handle(stream, self.device))
        return n - k

    def set_dispatch_params(self, params):
        """Installs the dispatcher network's parameters (RZ_MLP_PARAMS floats) for
        remove_auto; None removes them (remove_auto then uses the k < n/2 rule)."""
        if params is None:
            self._call("rz_dispatch_set_params", self._h, None)
            return
        import numpy as np

        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != 2660:
            raise ValueError("the dispatcher has 2660 parameters")
        self._call("rz_dispatch_set_params", self._h, p.ctypes.data_as(ctypes.c_void_p))

    def remove_auto(self, data, removals, n=None, k=None, flags=0, stream=None):
        """remove() with the method picked per list by the installed network (or
        the k < n/2 rule); returns (n - k, RZ_METHOD_* used).  Reads the choice
        back, so it synchronises the stream."""
        n, k, esz = self._args(data, removals, n, k)
        m = ctypes.c_int32(-1)
        self._call("rz_remove_auto", self._h, ctypes.c_void_p(data.data_ptr()), n, esz,
                   ctypes.c_void_p(removals.data_ptr()), k, flags, ctypes.byref(m), _stream_handle(stream, self.device))
        return n - k, m.value

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

    STAGE_NAMES = ("marks_done", "ready_raised", "first_slice_start", "last_slice_start", "ready_wait_end",
                   "lookback_end", "moves_end", "exchange_end", "exit", "mean_ready_wait", "mean_lookback",
                   "mean_moves", "mean_exchange", "mean_exit_tail", "mark_groups", "slices")

    def last_stages(self):
        """Stage timeline (us from the first workgroup's start) of the last
        fused-form call made with RZ_PROFILE (see rz_last_stages)."""
        out = (ctypes.c_float * 16)()
        self._call("rz_last_stages", self._h, out)
        return dict(zip(self.STAGE_NAMES, [round(v, 3) for v in out]))

    def last_stats(self):
        out = Stats()
        self._call("rz_last_stats", self._h, ctypes.byref(out))
        return out

    def check_errors(self, stream=None):
        """Waits for the stream and reads (and clears) the device error word.
        Returns 0, or raises RedzoneError(RZ_EDEVICE) carrying the RZ_ERR_*
        bits: the calls since the last check may have produced wrong arrays
        (the library has reset its scratch)."""
        flags = ctypes.c_uint32(0)
        st = self._lib.rz_check_errors(self._h, _stream_handle(stream, self.device), ctypes.byref(flags))
        if st != 0:
            raise RedzoneError(st, "rz_check_errors", flags.value)
        return 0

    def debug_set_waits(self, spin_cap=0, stall=False):
        """Test hook: in-kernel waits give up after spin_cap polls (0 = their own
        limits); with stall the Redzone waits never see their condition met."""
        self._call("rz_debug_set_waits", self._h, int(spin_cap), 1 if stall else 0)

    # ---- synthetic inputs --------------------------------------------------
    def gen_uniform_removals(self, removals, n, seed, stream=None):
        """Fills the int32 device tensor `removals` (k entries) with k distinct indices of [0, n)."""
        _check_tensor(removals, "removals", self.device)
        self._call("rz_gen_uniform_removals", self._h, ctypes.c_void_p(removals.data_ptr()), int(n),
                   removals.shape[0], int(seed), _stream_handle(stream, self.device))

    def gen_payload(self, data, seed, stream=None):
        _check_tensor(data, "data", self.device)
        self._call("rz_gen_payload", self._h, ctypes.c_void_p(data.data_ptr()), data.shape[0], elem_bytes_of(data),
                   int(seed), _stream_handle(stream, self.device))

    # ---- autotuner ---------------------------------------------------------
    @staticmethod
    def candidates(batch=False):
        lib = _native.load()
        cnt = ctypes.c_int32()
        lib.rz_tune_candidates(1 if batch else 0, None, 0, ctypes.byref(cnt))
        arr = (TileConfig * cnt.value)()
        lib.rz_tune_candidates(1 if batch else 0, arr, cnt.value, ctypes.byref(cnt))
        return list(arr)

    def tune(self, elem_bytes, n, k, candidates=None, reps=5, cold=True):
        """Times every candidate on a synthetic list, installs the fastest; returns
        (best, best_ms, [(cfg, ms), ...]) with ms = -1 for shapes the device refused."""
        cands = candidates or self.candidates(False)
        arr = (TileConfig * len(cands))(*cands)
        all_ms = (ctypes.c_float * len(cands))()
        best = TileConfig()
        best_ms = ctypes.c_float()
        self._call("rz_tune", self._h, elem_bytes, int(n), int(k), arr, len(cands), reps, 1 if cold else 0,
                   ctypes.byref(best), ctypes.byref(best_ms), all_ms)
        return best, best_ms.value, list(zip(cands, list(all_ms)))

    def tune_batch(self, elem_bytes, nlists, n_min, n_max, r_max, reps=3, candidates=None):
        cands = candidates or self.candidates(True)
        arr = (TileConfig * len(cands))(*cands)
        all_ms = (ctypes.c_float * len(cands))()
        best = TileConfig()
        best_ms = ctypes.c_float()
        self._call("rz_tune_batch", self._h, elem_bytes, nlists, int(n_min), int(n_max), float(r_max), reps, arr,
                   len(cands), ctypes.byref(best), ctypes.byref(best_ms), all_ms)
        return best, best_ms.value, list(zip(cands, list(all_ms)))

    def table(self, elem_bytes, log2k):
        out = TileConfig()
        self._call("rz_table_get", self._h, elem_bytes, log2k, ctypes.byref(out))
        return out

    def set_table(self, elem_bytes, log2k, cfg):
        self._call("rz_table_set", self._h, elem_bytes, log2k, ctypes.byref(cfg))

    def batch_config(self):
        out = TileConfig()
        self._call("rz_batch_config_get", self._h, ctypes.byref(out))
        return out

    def set_batch_config(self, cfg):
        self._call("rz_batch_config_set", self._h, ctypes.byref(cfg))

    def save_table(self, path):
        self._call("rz_table_save", self._h, os.fsencode(path))

    def load_table(self, path):
        self._call("rz_table_load", self._h, os.fsencode(path))

    def _descs(self, lists):
        """lists: iterable of (data, removals) device tensors -> ctypes array of ListDesc."""
        lists = list(lists)
        arr = (ListDesc * max(1, len(lists)))()
        for i, (data, rem) in enumerate(lists):
            n, k, esz = self._args(data, rem, None, None)
            arr[i] = ListDesc(ctypes.c_void_p(data.data_ptr()), ctypes.c_void_p(rem.data_ptr()), n, k, esz, 0)
        return arr, len(lists)

    def batch(self, lists, flags=0, config=None):
        """Plans a batch over [(data, removals), ...] or a ListDescs; returns a Batch."""
        if isinstance(lists, ListDescs):
            arr, cnt = lists.as_ctypes(), len(lists)
        else:
            arr, cnt = self._descs(lists)
        h = ctypes.c_void_p()
        cfg = ctypes.byref(config) if config is not None else None
        self._call("rz_batch_create", self._h, arr, cnt, flags, cfg, ctypes.byref(h))
        return Batch(self, h)

    def descs(self, lists):
        """ListDescs (the C ABI's rz_list_desc rows) of [(data, removals), ...],
        built once so that batches, feature extraction and per-list launches can
        share it."""
        import numpy as np
        import torch

        i32 = torch.int32
        rows = []
        for data, rem in lists:
            if not (data.is_cuda and rem.is_cuda) or rem.dtype is not i32:
                raise ValueError("lists must hold device tensors and int32 removals")
            if data.get_device() != self.device or rem.get_device() != self.device:
                raise ValueError(f"tensors must be on cuda:{self.device}")
            if not (data.is_contiguous() and rem.is_contiguous()):
                raise ValueError("tensors must be contiguous")
            rows.append((data.data_ptr(), rem.data_ptr(), data.shape[0], rem.shape[0], elem_bytes_of(data), 0))
        return ListDescs(np.array(rows, dtype=LIST_DESC_DTYPE) if rows else np.zeros(0, LIST_DESC_DTYPE))

    def remove_batched(self, lists, flags=0, stream=None):
        """One-shot batched removal over [(data, removals), ...] or a ListDescs."""
        if isinstance(lists, ListDescs):
            arr, cnt = lists.as_ctypes(), len(lists)
        else:
            arr, cnt = self._descs(lists)
        self._call("rz_remove_batched", self._h, arr, cnt, flags, _stream_handle(stream, self.device))


# rz_list_desc as a numpy record (40 bytes, the C layout)
def _list_desc_dtype():
    import numpy as np

    return np.dtype([("data", "<u8"), ("removals", "<u8"), ("n", "<i8"), ("k", "<i8"), ("elem_bytes", "<u4"),
                     ("method", "<u4")])


LIST_DESC_DTYPE = _list_desc_dtype()
assert LIST_DESC_DTYPE.itemsize == ctypes.sizeof(ListDesc)


class ListDescs:
    """Host array of rz_list_desc rows (see Context.descs); subset(idx) selects rows."""

    def __init__(self, rows):
        import numpy as np

        self.rows = np.ascontiguousarray(rows, dtype=LIST_DESC_DTYPE)

    def __len__(self):
        return len(self.rows)

    def subset(self, idx):
        import numpy as np

        return ListDescs(self.rows[np.asarray(idx, np.int64)])

    def as_ctypes(self):
        return self.rows.ctypes.data_as(ctypes.POINTER(ListDesc))


class Batch:
    """A planned batch of lists (rz_batch): launch it, or capture it in a hipGraph and replay."""

    def __init__(self, ctx, handle):
        self._ctx = ctx
        self._lib = ctx._lib
        self._h = handle

    def launch(self, stream=None):
        self._ctx._call("rz_batch_launch", self._h, _stream_handle(stream, self._ctx.device))

    def capture(self, stream=None):
        self._ctx._call("rz_batch_capture", self._h, _stream_handle(stream, self._ctx.device))

    def info(self):
        out = BatchInfo()
        self._ctx._call("rz_batch_get_info", self._h, ctypes.byref(out))
        return {"nitems": out.nitems, "bitmap_bytes": out.bitmap_bytes, "grid": out.grid,
                "lists_active": out.lists_active}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rz_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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
