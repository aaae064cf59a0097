"""rz_remove_noalloc, the allocation-free entry point (scratch in one fixed
device array), through the C ABI: its survivors must be array-identical to
those of rz_remove (the entry point it varies) and to the canonical oracle, on
two calls in a row with the second one smaller, for every item size; its
limit and flag checks return before anything runs; the caller's error word
receives the device checks' bits."""
import ctypes

import numpy as np
import pytest
import torch

import redzone_amd as rz
from redzone_amd import _native
from redzone_amd import workloads as W
from tests.parity import announce, expected_canonical, first_mismatch

pytestmark = pytest.mark.gpu


def _noalloc(A, R, flags=0, err=None, k=None):
    lib = _native.load()
    k = R.shape[0] if k is None else k
    return lib.rz_remove_noalloc(ctypes.c_void_p(A.data_ptr()), A.shape[0], rz.elem_bytes_of(A),
                                 ctypes.c_void_p(R.data_ptr()), k, flags,
                                 ctypes.c_void_p(err.data_ptr()) if err is not None else None,
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def _both(ctx, oracle, A, R, flags=0):
    """Survivors of rz_remove_noalloc and of rz_remove on the same inputs; both
    must equal the canonical oracle's."""
    n, k = A.shape[0], R.shape[0]
    err = torch.zeros(1, dtype=torch.int32, device="cuda")
    a1, r1 = A.cuda(), R.cuda()
    assert _noalloc(a1, r1, flags, err) == 0
    a2, r2 = A.cuda(), R.cuda()
    ctx.remove(a2, r2)
    torch.cuda.synchronize()
    assert int(err.item()) == 0, f"device error bits {int(err.item()):#x}"
    assert ctx.check_errors() == 0
    got, ref = a1.cpu().numpy(), a2.cpu().numpy()
    assert first_mismatch(got[: n - k], ref[: n - k]) < 0
    want = expected_canonical(oracle, A.numpy(), R.numpy())
    assert first_mismatch(got[: n - k], want[: n - k]) < 0
    return r1


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_noalloc_matches_rz_remove_two_calls(ctx, oracle, esz):
    # the first call at the limit (mark kernel + fused kernel), the second
    # smaller (one-launch form) on the same array
    for n, k, pattern in [((1 << 23) + 3, _native.RZ_NOALLOC_MAX_K if esz == 4 else 1 << 21, "uniform"),
                          ((1 << 20) + 5, 3000, "tail_heavy")]:
        announce("noalloc", esz, n, k, pattern)
        A = W.payload(esz, n, seed=300 + esz)
        R = W.removals(pattern, n, k, seed=301 + esz)
        _both(ctx, oracle, A, R)


def test_noalloc_flags_and_shapes(ctx, oracle):
    n = (1 << 20) + 1
    for k, pattern in [(1, "uniform"), (200_000, "clustered"), (600_000, "tail_heavy"), (n - 1, "uniform")]:
        A = W.unique_u32_payload(n, seed=310)
        R = W.removals(pattern, n, k, seed=311)
        announce("noalloc sorted", k, pattern)
        _both(ctx, oracle, A, torch.sort(R).values.contiguous(), flags=rz.RZ_SORTED)
        announce("noalloc preserve", k, pattern)
        r1 = _both(ctx, oracle, A, R, flags=rz.RZ_PRESERVE_REMOVALS)
        assert torch.equal(r1.cpu(), R), "the removal list is left as it was"
        announce("noalloc any pairing", k, pattern)
        a, r = A.cuda(), R.cuda()
        assert _noalloc(a, r, rz.RZ_ANY_PAIRING) == 0
        torch.cuda.synchronize()
        bad, fb = oracle.check(A.numpy(), a.cpu().numpy(), R.numpy(), mode=0)
        assert bad == 0, (k, pattern, bad, fb)


def test_noalloc_limits(ctx):
    n = (1 << 23) + 3
    A = W.unique_u32_payload(n, seed=320).cuda()
    R = W.uniform_removals(n, _native.RZ_NOALLOC_MAX_K + 1, seed=321).cuda()
    before = A.clone()
    announce("noalloc limits")
    assert _noalloc(A, R) == _native.RZ_ETOOBIG
    for bad_flag in (rz.RZ_VALIDATE, rz.RZ_STATS, rz.RZ_PROFILE, rz.RZ_COOPERATIVE):
        assert _noalloc(A, R[:100], bad_flag) == _native.RZ_EINVAL
    assert _noalloc(A, R, k=n + 1) == _native.RZ_EINVAL
    torch.cuda.synchronize()
    assert torch.equal(A, before), "a rejected call runs nothing"
    assert _noalloc(A, R[:0]) == 0 and _noalloc(A[:5], R[:5]) == 0  # k == 0 and k == n return at once


def test_noalloc_reports_orphan_mismatch(ctx, oracle):
    """A list that breaks the contract (the last index twice): the call ORs
    RZ_ERR_ORPHAN_MISMATCH into the caller's word; the next valid call is
    exact and reports nothing."""
    n, k = 4096, 64
    A = W.unique_u32_payload(n, seed=330).cuda()
    head = W.uniform_removals(n - k, k - 2, seed=331)
    R = torch.cat([torch.tensor([n - 1, n - 1], dtype=torch.int32), head]).cuda()
    err = torch.zeros(1, dtype=torch.int32, device="cuda")
    announce("noalloc mismatch")
    assert _noalloc(A, R, 0, err) == 0
    torch.cuda.synchronize()
    assert int(err.item()) & rz.RZ_ERR_ORPHAN_MISMATCH
    A2 = W.unique_u32_payload(70_000, seed=332)
    R2 = W.uniform_removals(70_000, 700, seed=333)
    _both(ctx, oracle, A2, R2)
