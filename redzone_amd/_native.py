"""ctypes binding of librz_amd.so (the C ABI declared in include/redzone.h).

The product path is the HIP library; if it is missing this module raises
instead of falling back to anything else.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librz_amd.so")

# status codes (rz_status)
RZ_OK, RZ_EINVAL, RZ_ERANGE, RZ_EDUP, RZ_ECOOP, RZ_EHIP, RZ_ENOMEM, RZ_ENOTUNED, RZ_EDEVICE, RZ_ETOOBIG = range(10)
RZ_NOALLOC_MAX_K = 1 << 22
# bits of the device error word (RZ_ERR_*)
RZ_ERR_ORPHAN_MISMATCH = 0x1
RZ_ERR_WAIT_TIMEOUT = 0x2
RZ_ERR_LOOKBACK_TIMEOUT = 0x4
RZ_ERR_SIZES = 0x8
# flags
RZ_PRESERVE_REMOVALS = 0x1
RZ_VALIDATE = 0x2
RZ_COOPERATIVE = 0x4
RZ_STATS = 0x8
RZ_PROFILE = 0x10
RZ_SORTED = 0x20
RZ_ANY_PAIRING = 0x40
# dispatcher methods (RZ_METHOD_*)
RZ_METHOD_REDZONE, RZ_METHOD_ORDERED, RZ_METHOD_DENSE, RZ_METHOD_PARTITION = range(4)


class TileConfig(ctypes.Structure):
    _fields_ = [
        ("wg_size", ctypes.c_int32),
        ("items", ctypes.c_int32),
        ("max_slices", ctypes.c_int32),
        ("variant", ctypes.c_int32),
        ("slice_steps", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
    ]

    def __repr__(self):
        return (f"TileConfig(wg_size={self.wg_size}, items={self.items}, max_slices={self.max_slices}, "
                f"variant={self.variant}, slice_steps={self.slice_steps}, chunk={self.chunk})")

    def as_tuple(self):
        return (self.wg_size, self.items, self.max_slices, self.variant, self.slice_steps, self.chunk)


class ListDesc(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("removals", ctypes.c_void_p),
        ("n", ctypes.c_int64),
        ("k", ctypes.c_int64),
        ("elem_bytes", ctypes.c_uint32),
        ("method", ctypes.c_uint32),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("nitems", ctypes.c_int64),
        ("bitmap_bytes", ctypes.c_int64),
        ("grid", ctypes.c_int32),
        ("lists_active", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("k", ctypes.c_int64),
        ("tail_removals", ctypes.c_int64),
        ("orphan_pairs", ctypes.c_int64),
        ("slices", ctypes.c_int64),
        ("error_flags", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class RedzoneError(RuntimeError):
    """A non-OK rz_status; error_flags holds the device error word's RZ_ERR_* bits for RZ_EDEVICE."""

    def __init__(self, status, what="", error_flags=0):
        self.status = status
        self.error_flags = error_flags
        msg = _lib.rz_status_string(status).decode() if _lib is not None else str(status)
        if error_flags:
            msg += f", error bits 0x{error_flags:x}"
        super().__init__(f"{what}: {msg} (status {status})" if what else f"{msg} (status {status})")


_lib = None

# (name, restype, argtypes) of every exported function
_vp, _i64, _u32, _i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint32, ctypes.c_int32
SIGNATURES = [
    ("rz_ctx_create", _i32, [ctypes.POINTER(_vp), _i32]),
    ("rz_ctx_destroy", _i32, [_vp]),
    ("rz_status_string", ctypes.c_char_p, [_i32]),
    ("rz_device_info", _i32, [_vp, ctypes.POINTER(_i32), ctypes.c_char_p, _i32]),
    ("rz_remove", _i32, [_vp, _vp, _i64, _u32, _vp, _i64, _u32, _vp]),
    ("rz_remove_u32", _i32, [_vp, _vp, _i64, _vp, _i64, _u32, _vp]),
    ("rz_remove_u64", _i32, [_vp, _vp, _i64, _vp, _i64, _u32, _vp]),
    ("rz_remove_rec16", _i32, [_vp, _vp, _i64, _vp, _i64, _u32, _vp]),
    ("rz_remove_config", _i32, [_vp, _vp, _i64, _u32, _vp, _i64, _u32, ctypes.POINTER(TileConfig), _vp]),
    ("rz_last_stats", _i32, [_vp, ctypes.POINTER(Stats)]),
    ("rz_last_profile", _i32, [_vp, ctypes.POINTER(ctypes.c_float)]),
    ("rz_partition_on", _i32, [_vp, _vp, _i64, _u32, _vp, _i64, _u32, _vp]),
    ("rz_batch_create", _i32, [_vp, ctypes.POINTER(ListDesc), _i32, _u32, ctypes.POINTER(TileConfig),
                               ctypes.POINTER(_vp)]),
    ("rz_batch_launch", _i32, [_vp, _vp]),
    ("rz_batch_capture", _i32, [_vp, _vp]),
    ("rz_batch_get_info", _i32, [_vp, ctypes.POINTER(BatchInfo)]),
    ("rz_batch_destroy", _i32, [_vp]),
    ("rz_remove_batched", _i32, [_vp, ctypes.POINTER(ListDesc), _i32, _u32, _vp]),
    ("rz_gen_uniform_removals", _i32, [_vp, _vp, _i64, _i64, ctypes.c_uint64, _vp]),
    ("rz_gen_payload", _i32, [_vp, _vp, _i64, _u32, ctypes.c_uint64, _vp]),
    ("rz_tune_candidates", _i32, [_i32, ctypes.POINTER(TileConfig), _i32, ctypes.POINTER(_i32)]),
    ("rz_tune", _i32, [_vp, _u32, _i64, _i64, ctypes.POINTER(TileConfig), _i32, _i32, _i32,
                       ctypes.POINTER(TileConfig), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    ("rz_tune_batch", _i32, [_vp, _u32, _i32, _i64, _i64, ctypes.c_double, _i32, ctypes.POINTER(TileConfig), _i32,
                             ctypes.POINTER(TileConfig), ctypes.POINTER(ctypes.c_float),
                             ctypes.POINTER(ctypes.c_float)]),
    ("rz_table_get", _i32, [_vp, _u32, _i32, ctypes.POINTER(TileConfig)]),
    ("rz_table_set", _i32, [_vp, _u32, _i32, ctypes.POINTER(TileConfig)]),
    ("rz_batch_config_get", _i32, [_vp, ctypes.POINTER(TileConfig)]),
    ("rz_batch_config_set", _i32, [_vp, ctypes.POINTER(TileConfig)]),
    ("rz_table_save", _i32, [_vp, ctypes.c_char_p]),
    ("rz_table_load", _i32, [_vp, ctypes.c_char_p]),
    ("rz_features", _i32, [_vp, _vp, _i32, _vp, _vp]),
    ("rz_mlp_forward", _i32, [_vp, _vp, _vp, _i32, _vp, _vp, _vp]),
    ("rz_mlp_grad", _i32, [_vp, _vp, _vp, _vp, _vp, _i32, ctypes.c_float, _vp, _vp, _vp, _vp]),
    ("rz_mlp_adam", _i32, [_vp, _vp, _vp, _vp, _vp, _i32, ctypes.c_float, _vp]),
    ("rz_remove_premarked", _i32, [_vp, _vp, _i64, _u32, _vp, _i64, _vp, _u32, ctypes.POINTER(TileConfig), _vp]),
    ("rz_mark_removals", _i32, [_vp, _vp, _i64, _i64, _vp, _vp]),
    ("rz_add_remove", _i32, [_vp, _vp, _i64, _i64, _u32, _vp, _i64, _vp, _i64, _u32, _vp]),
    ("rz_gather_preserve", _i32, [_vp, _vp, _i64, _u32, _vp, _i64, _u32, _vp]),
    ("rz_dispatch_method_config", _i32, [_i32, _u32, ctypes.POINTER(TileConfig)]),
    ("rz_dispatch_set_params", _i32, [_vp, _vp]),
    ("rz_remove_auto", _i32, [_vp, _vp, _i64, _u32, _vp, _i64, _u32, ctypes.POINTER(_i32), _vp]),
    ("rz_dispatch_choose", _i32, [_vp, _vp, ctypes.POINTER(ListDesc), _i32, _vp]),
    ("rz_dispatch_collect", _i32, [_vp, ctypes.POINTER(_i32), _i32]),
    ("rz_remove_auto_batched", _i32, [_vp, _vp, ctypes.POINTER(ListDesc), _i32, _i64, _u32, ctypes.POINTER(_i32), _vp]),
    ("rz_check_errors", _i32, [_vp, _vp, ctypes.POINTER(_u32)]),
    ("rz_debug_set_waits", _i32, [_vp, _u32, _u32]),
    ("rz_remove_noalloc", _i32, [_vp, _i64, _u32, _vp, _i64, _u32, _vp, _vp]),
    ("rz_remove_sized", _i32, [_vp, _vp, _i64, _u32, _vp, _i64, _vp, _vp, _u32, _vp]),
    ("rz_last_stages", _i32, [_vp, ctypes.POINTER(ctypes.c_float)]),
]

TUNED_TABLE = os.path.join(_HERE, "tuned_gfx950.txt")


def _hip_runtimes_loaded():
    """Paths of the HIP runtime libraries mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            return sorted({line.split()[-1] for line in f if "libamdhip64" in line})
    except OSError:
        return []


def load():
    """Load librz_amd.so; raise if it has not been built.

    torch ships its own HIP runtime with the same soname as the system one
    (libamdhip64.so.7).  torch is imported first so that the library binds to
    the runtime torch already loaded: one runtime means torch's streams and
    device memory are this library's streams and memory.  Two runtimes in one
    process would not share streams, so that case raises."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        runtimes = _hip_runtimes_loaded()
        if len(runtimes) > 1:
            raise RuntimeError(f"two HIP runtimes are loaded ({runtimes}); import torch before redzone_amd")
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
