"""rz_remove_sized: n and k read from device memory, as the reference's launch
wrapper reads them (TestGPUWide(A, sizeA, R, sizeR, ...) dereferences sizeA
and sizeR inside the kernel, redzone.hxx:457-465).

Every result is compared with the canonical oracle on (A[:n], R[:k]) for the
device values; the items past n and the removal entries past k (filled with
values no valid list holds) must not be read or written.  A captured graph of
one call removes whatever k the device holds at each replay.  Device sizes
that break their bounds are refused inside the kernel (every workgroup
leaves at once) and reported through the error word."""
import numpy as np
import pytest
import torch

import redzone_amd as rz
from redzone_amd import workloads as W
from tests.parity import announce, expected_canonical, first_mismatch

pytestmark = pytest.mark.gpu


def _sizes(n, k):
    return (torch.tensor([n], dtype=torch.int32, device="cuda"),
            torch.tensor([k], dtype=torch.int32, device="cuda"))


def _check(oracle, A0, R0, n, k, got, any_pairing=False):
    """got: the whole n_cap buffer after the call; A0, R0: the host inputs
    (n_cap items, k_cap entries)."""
    A_np, R_np = A0.numpy(), np.ascontiguousarray(R0.numpy()[:k])
    if any_pairing:
        bad, fb = oracle.check(A_np[:n], got[:n], R_np, mode=0)
        assert bad == 0, (n, k, bad, fb)
    else:
        want = expected_canonical(oracle, A_np[:n], R_np)
        m = first_mismatch(got[: n - k], want[: n - k])
        assert m < 0, f"first mismatch at {m} (n={n}, k={k})"
    assert np.array_equal(got[n:], A_np[n:]), "an item past n was written"


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_sized_equals_canonical(ctx, oracle, esz):
    """Device sizes equal to the buffers: the same array as rz_remove gives."""
    for n, k in [(1000, 10), (1 << 16, 1 << 15), ((1 << 20) + 7, 1 << 14), ((1 << 22) + 3, 40_000),
                 ((1 << 22) + 3, 600_000)]:
        A = W.payload(esz, n, seed=600 + esz)
        R = W.uniform_removals(n, k, seed=601)
        announce("sized", esz, n, k)
        dA, dR = A.cuda(), R.cuda()
        d_n, d_k = _sizes(n, k)
        ctx.remove_sized(dA, dR, d_n, d_k)
        torch.cuda.synchronize()
        assert ctx.check_errors() == 0
        _check(oracle, A, R, n, k, dA.cpu().numpy())


def test_sized_below_caps(ctx, oracle):
    """n below n_cap and k below k_cap: only A[:n] and R[:k] take part (the
    entries past k hold indices no valid list has); k far below k_cap takes
    the shape the device k gives (more slices than k_cap's slice length would)."""
    n_cap, k_cap = (1 << 22) + 11, 1 << 20
    A = W.unique_u32_payload(n_cap, seed=610)
    for n, k, pattern in [(n_cap - 5, k_cap, "uniform"), ((1 << 21) + 1, 300_000, "clustered"),
                          (1 << 20, 3_000, "tail_heavy"), (5000, 1, "uniform"), (n_cap, 65_536, "uniform")]:
        R = torch.full((k_cap,), -7, dtype=torch.int32)
        R[:k] = W.removals(pattern, n, k, seed=611 + k)
        R[k + 1::2] = n_cap + 9
        announce("sized below caps", n, k, pattern)
        dA, dR = A.cuda(), R.cuda()
        d_n, d_k = _sizes(n, k)
        ctx.remove_sized(dA, dR, d_n, d_k, n_cap=n_cap, k_cap=k_cap)
        torch.cuda.synchronize()
        assert ctx.check_errors() == 0
        _check(oracle, A, R, n, k, dA.cpu().numpy())
        assert torch.equal(dR[k:].cpu(), R[k:]), "an entry past k was written"


@pytest.mark.parametrize("flags", [rz.RZ_ANY_PAIRING, rz.RZ_SORTED, rz.RZ_STATS | rz.RZ_PRESERVE_REMOVALS])
def test_sized_flags(ctx, oracle, flags):
    n, k = (1 << 21) + 3, 200_000
    A = W.payload(8, n, seed=620)
    R = W.uniform_removals(n, k, seed=621)
    if flags & rz.RZ_SORTED:
        R = torch.sort(R).values
    announce("sized flags", flags)
    dA, dR = A.cuda(), R.cuda()
    d_n, d_k = _sizes(n, k)
    ctx.remove_sized(dA, dR, d_n, d_k, flags=flags)
    torch.cuda.synchronize()
    _check(oracle, A, R, n, k, dA.cpu().numpy(), any_pairing=bool(flags & rz.RZ_ANY_PAIRING))
    if flags & rz.RZ_STATS:
        st = ctx.last_stats()
        assert st.tail_removals == int((R.numpy() >= n - k).sum()) and st.error_flags == 0
        assert torch.equal(dR.cpu(), R), "RZ_PRESERVE_REMOVALS: the list changed"
    assert ctx.check_errors() == 0


def test_sized_nothing_to_move(ctx):
    """Device k = 0 and k = n: nothing changes and nothing is reported."""
    n = 4096
    A = W.unique_u32_payload(n, seed=630)
    R = W.uniform_removals(n, n, seed=631)
    dA, dR = A.cuda(), R.cuda()
    for k in (0, n):
        announce("sized nothing", k)
        d_n, d_k = _sizes(n, k)
        ctx.remove_sized(dA, dR, d_n, d_k)
        torch.cuda.synchronize()
        assert ctx.check_errors() == 0
        assert torch.equal(dA.cpu(), A)


def test_sized_bad_device_sizes(ctx, oracle):
    """Device sizes outside their bounds: every workgroup leaves at once,
    nothing is written, RZ_ERR_SIZES reaches the caller, and the next call on
    the context is exact."""
    n_cap, k_cap = 1 << 16, 1 << 12
    A = W.unique_u32_payload(n_cap, seed=640)
    R = W.uniform_removals(n_cap, k_cap, seed=641)
    dA, dR = A.cuda(), R.cuda()
    for n, k in [(n_cap, k_cap + 1), (100, 200), (n_cap + 1, 10), (n_cap, -1), (-4, 0)]:
        announce("sized bad sizes", n, k)
        d_n, d_k = _sizes(n, k)
        ctx.remove_sized(dA, dR, d_n, d_k)
        with pytest.raises(rz.RedzoneError) as e:
            ctx.check_errors()
        assert e.value.status == rz.RZ_EDEVICE and e.value.error_flags == rz.RZ_ERR_SIZES
        assert torch.equal(dA.cpu(), A) and torch.equal(dR.cpu(), R)
    d_n, d_k = _sizes(n_cap, k_cap)
    ctx.remove_sized(dA, dR, d_n, d_k)
    torch.cuda.synchronize()
    assert ctx.check_errors() == 0
    _check(oracle, A, R, n_cap, k_cap, dA.cpu().numpy())


def test_sized_host_refusals(ctx):
    """What the host can see is refused before any kernel: RZ_VALIDATE (k is on
    the device), and a k_cap no fused shape fits."""
    A = torch.zeros(1 << 10, dtype=torch.int32, device="cuda")
    R = torch.zeros(16, dtype=torch.int32, device="cuda")
    d_n, d_k = _sizes(1 << 10, 16)
    with pytest.raises(rz.RedzoneError) as e:
        ctx.remove_sized(A, R, d_n, d_k, flags=rz.RZ_VALIDATE)
    assert e.value.status == 1
    big_k = 60_000_000  # above 4096 slices of the longest slice that fits LDS
    A = torch.empty(1 << 26, dtype=torch.int32, device="cuda")
    R = torch.empty(big_k, dtype=torch.int32, device="cuda")
    with pytest.raises(rz.RedzoneError) as e:
        ctx.remove_sized(A, R, d_n, d_k)
    assert e.value.status == 1
    del A, R
    torch.cuda.empty_cache()


def test_sized_graph_replay_device_k(ctx, oracle):
    """One captured call; before each replay the removal count is changed on
    the device (an eager fill, no host-side size) together with restoring the
    inputs, and each replay equals the oracle for that k."""
    n, k_cap = (1 << 21) + 5, 1 << 18
    A = W.unique_u32_payload(n, seed=650)
    R = W.uniform_removals(n, k_cap, seed=651)  # every prefix is a valid list
    dA, dR = A.cuda(), R.cuda()
    dA0, dR0 = A.cuda(), R.cuda()
    d_n, d_k = _sizes(n, k_cap)
    ctx.remove_sized(dA, dR, d_n, d_k)  # eager call first: grows the scratch
    torch.cuda.synchronize()
    assert ctx.check_errors() == 0
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.remove_sized(dA, dR, d_n, d_k)
    for k in (k_cap, k_cap // 3, 1, 0, 99_999, k_cap):
        announce("sized graph replay", k)
        dA.copy_(dA0)
        dR.copy_(dR0)
        d_k.fill_(k)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert ctx.check_errors() == 0
        _check(oracle, A, R, n, k, dA.cpu().numpy())
