"""A native C++ host program (tests/native/abi_host.cpp, built by
__graft_entry__.build()) drives the library through include/redzone.h
only: rz_remove_u32, rz_partition_on, rz_remove_batched and the status
codes, checked with the invariants of the reference's harness
(DataStruct.hxx:154-191) on its test list A[i] = i."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "native", "abi_host")


def _binary():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    return BIN


def test_native_host_links_the_library():
    """The program loads librz_amd.so from the tree and reaches the ABI; without
    a GPU rz_ctx_create fails with a status code (exit 77), never a crash."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present (the GPU test runs the program)")
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 77, (r.returncode, r.stdout, r.stderr)
    assert "rz_ctx_create" in r.stdout


@pytest.mark.gpu
def test_native_host_passes_on_the_gpu():
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASS" in r.stdout, (r.returncode, r.stdout, r.stderr)
