"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker or as the timed CPU
baseline, never as a code path of the product.

Every function takes numpy arrays (C-contiguous) and works in place where the
C function does.  Element arrays may have any dtype; they are handed over as
raw bytes with an element size `esz`.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OK, EINVAL, ERANGE, EDUP, ENOMEM = 0, -1, -2, -3, -4


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i64, u32, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint32, ctypes.c_int
        L.rzo_canonical.argtypes = [vp, i64, u32, vp, i64, i32]
        L.rzo_canonical.restype = i32
        L.rzo_count_tail.argtypes = [vp, i64, i64]
        L.rzo_count_tail.restype = i64
        L.rzo_redzone_cpu.argtypes = [vp, i64, u32, vp, i64, i32, vp]
        L.rzo_redzone_cpu.restype = i32
        L.rzo_remove_stable.argtypes = [vp, i64, u32, vp, i64, i32, vp]
        L.rzo_remove_stable.restype = i32
        L.rzo_check.argtypes = [vp, vp, i64, u32, vp, i64, i32, vp]
        L.rzo_check.restype = i32
        L.rzo_check_multiset.argtypes = [vp, vp, i64, u32, vp, i64]
        L.rzo_check_multiset.restype = i32
        _lib = L
    return _lib


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def _esz(a, n):
    return a.nbytes // n if n else a.itemsize


def canonical(data, R, mutation=0):
    """Sequential canonical Redzone, in place on `data` (shape (n, ...))."""
    n = data.shape[0]
    R = np.ascontiguousarray(R, dtype=np.int32)
    return lib().rzo_canonical(_ptr(data), n, _esz(data, n), _ptr(R), R.shape[0], mutation)


def count_tail(R, n):
    R = np.ascontiguousarray(R, dtype=np.int32)
    return lib().rzo_count_tail(_ptr(R), n, R.shape[0])


def redzone_cpu(data, R, threads=0):
    """Redzone_CPU restatement; clobbers R like the reference.  Returns
    (status, [total, phase1, phase2, phase3] in ns)."""
    n = data.shape[0]
    assert R.dtype == np.int32 and R.flags["C_CONTIGUOUS"]
    times = np.zeros(4, dtype=np.int64)
    st = lib().rzo_redzone_cpu(_ptr(data), n, _esz(data, n), _ptr(R), R.shape[0], threads, _ptr(times))
    return st, times


def remove_stable(data, R, threads=1):
    n = data.shape[0]
    R = np.ascontiguousarray(R, dtype=np.int32)
    t = np.zeros(1, dtype=np.int64)
    st = lib().rzo_remove_stable(_ptr(data), n, _esz(data, n), _ptr(R), R.shape[0], threads, _ptr(t))
    return st, int(t[0])


def check(orig, out, R, mode=0):
    """Invariant bitmask (0 = pass); mode 0 = Redzone, mode 1 = stable order."""
    n = orig.shape[0]
    R = np.ascontiguousarray(R, dtype=np.int32)
    fb = np.zeros(1, dtype=np.int64)
    bad = lib().rzo_check(_ptr(orig), _ptr(out), n, _esz(orig, n), _ptr(R), R.shape[0], mode, _ptr(fb))
    return bad, int(fb[0])


def check_multiset(orig, out, R):
    n = orig.shape[0]
    R = np.ascontiguousarray(R, dtype=np.int32)
    return lib().rzo_check_multiset(_ptr(orig), _ptr(out), n, _esz(orig, n), _ptr(R), R.shape[0])
