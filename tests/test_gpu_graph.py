"""Single-list removals captured in a hipGraph and replayed.

The fused forms keep their call epoch on the device (the last workgroup of a
call advances it), and rz_remove_noalloc clears its bitmap words and counters
with a prologue kernel, not a memset, so a captured call can be replayed.
Each test captures a call once, then restores the inputs with eager work and
a synchronisation before every replay and checks every replay against the
canonical oracle and the device error word."""
import ctypes

import numpy as np
import pytest
import torch

import redzone_amd as rz
from redzone_amd import _native
from redzone_amd import workloads as W
from tests.parity import announce, expected_canonical, first_mismatch

pytestmark = pytest.mark.gpu


def _noalloc(A, R, err):
    lib = _native.load()
    st = lib.rz_remove_noalloc(ctypes.c_void_p(A.data_ptr()), A.shape[0], rz.elem_bytes_of(A),
                               ctypes.c_void_p(R.data_ptr()), R.shape[0], 0, ctypes.c_void_p(err.data_ptr()),
                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, st


def _replay_and_check(oracle, g, A0, R0, dA, dR, want, err=None, ctx=None, reps=4):
    n, k = A0.shape[0], R0.shape[0]
    dA0, dR0 = A0.cuda(), R0.cuda()
    for rep in range(reps):
        dA.copy_(dA0)
        dR.copy_(dR0)
        if err is not None:
            err.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = dA.cpu().numpy()
        m = first_mismatch(got[: n - k], want[: n - k])
        assert m < 0, f"replay {rep}: first mismatch at {m}"
        if err is not None:
            assert int(err.item()) == 0, f"replay {rep}: error bits {int(err.item()):#x}"
        if ctx is not None:
            assert ctx.check_errors() == 0


@pytest.mark.parametrize("esz", [4, 16])
@pytest.mark.parametrize("k", [40_000, 600_000])  # the one-launch path, and the mark + fused kernel path
def test_noalloc_graph_replay(oracle, esz, k):
    n = (1 << 22) + 3
    A0 = W.payload(esz, n, seed=400 + esz)
    R0 = W.uniform_removals(n, k, seed=401)
    want = expected_canonical(oracle, A0.numpy(), R0.numpy())
    dA, dR = A0.cuda(), R0.cuda()
    err = torch.zeros(1, dtype=torch.int32, device="cuda")
    announce("noalloc graph", esz, k)
    _noalloc(dA, dR, err)  # eager call first
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        _noalloc(dA, dR, err)
    _replay_and_check(oracle, g, A0, R0, dA, dR, want, err=err)


@pytest.mark.parametrize("k", [40_000, 600_000])
def test_rz_remove_graph_replay(oracle, k):
    """rz_remove with the tuned table (the one-launch fused form at 40,000
    removals, mark + fused kernel at 600,000) on a context whose scratch was
    grown by an eager call of the same size before the capture."""
    n = (1 << 22) + 3
    ctx = rz.Context(0)
    try:
        A0 = W.unique_u32_payload(n, seed=410)
        R0 = W.uniform_removals(n, k, seed=411)
        want = expected_canonical(oracle, A0.numpy(), R0.numpy())
        dA, dR = A0.cuda(), R0.cuda()
        announce("rz_remove graph", k, ctx.table(4, k.bit_length() - 1).as_tuple())
        ctx.remove(dA, dR)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            ctx.remove(dA, dR)
        _replay_and_check(oracle, g, A0, R0, dA, dR, want, ctx=ctx)
    finally:
        ctx.close()
