"""The C-ABI library loads and exports exactly what include/redzone.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for fname in os.listdir(os.path.join(ROOT, "include")):
        if not fname.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fname)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        # device-side inline helpers (redzone_device.h) are not library exports
        text = re.sub(r"__device__[^{]*\{.*?\n\}", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*", "", text)
        names.update(re.findall(r"\b(rz_[a-z0-9_]+)\s*\(", text))
    return names


def test_device_header_declares_the_producer_helper():
    text = open(os.path.join(ROOT, "include", "redzone_device.h")).read()
    assert "__device__ inline void rz_mark_removal(" in text


def test_library_exports_every_declared_symbol():
    from redzone_amd import _native

    lib = _native.load()
    declared = _declared_functions()
    assert "rz_remove" in declared and "rz_partition_on" in declared
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"
    bound = {n for n, _, _ in _native.SIGNATURES}
    assert declared == bound, f"binding and header differ: {declared ^ bound}"


def test_status_strings():
    from redzone_amd import _native

    lib = _native.load()
    for s in range(10):
        assert lib.rz_status_string(s) and lib.rz_status_string(s) != b"unknown status"
    assert lib.rz_status_string(0) == b"ok"


def test_context_without_gpu_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from redzone_amd import _native

    lib = _native.load()
    h = ctypes.c_void_p()
    assert lib.rz_ctx_create(ctypes.byref(h), 0) != 0
    assert not h.value


def test_null_context_rejected():
    from redzone_amd import _native

    lib = _native.load()
    assert lib.rz_remove(None, None, 10, 4, None, 2, 0, None) == _native.RZ_EINVAL
    assert lib.rz_partition_on(None, None, 10, 4, None, 2, 0, None) == _native.RZ_EINVAL


def test_library_is_built_for_gfx950():
    from redzone_amd import _native

    blob = open(_native.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_dispatch_method_configs():
    """The dispatcher's fixed launch shapes come from the library (no device needed)."""
    import ctypes

    from redzone_amd import _native
    from redzone_amd.dispatch import method_config

    lib = _native.load()
    out = _native.TileConfig()
    assert lib.rz_dispatch_method_config(1, 4, ctypes.byref(out)) == 0
    assert out.as_tuple() == (256, 4, 4096, 3, 0, 1)
    for esz, items in ((4, 8), (8, 8), (16, 4)):
        assert lib.rz_dispatch_method_config(2, esz, ctypes.byref(out)) == 0
        assert out.as_tuple() == (1024, items, 1024, 5, 11, 2)
        assert method_config(2, esz).as_tuple() == out.as_tuple()
    assert method_config(0, 4) is None and method_config(3, 4) is None
    assert lib.rz_dispatch_method_config(0, 4, ctypes.byref(out)) == _native.RZ_EINVAL
    assert lib.rz_dispatch_method_config(3, 4, ctypes.byref(out)) == _native.RZ_EINVAL
    assert lib.rz_dispatch_method_config(2, 12, ctypes.byref(out)) == _native.RZ_EINVAL
