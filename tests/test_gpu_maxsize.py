"""The largest list the boundary accepts, n = 2^31 - 1 (the reference's sizes
are `int`, redzone.hxx:627): the fused Redzone forms, the dense form and the
O(n) partition array-exact against the oracle, and the one-pass preserve-list
gather against torch indexing on the device."""
import pytest
import torch

import redzone_amd as rz
from tests.parity import announce, expected_canonical, first_mismatch

pytestmark = pytest.mark.gpu

N_MAX = (1 << 31) - 1


@pytest.fixture(scope="module")
def big(ctx):
    free, _ = torch.cuda.mem_get_info()
    if free < (48 << 30):
        pytest.skip("needs 48 GiB of free device memory")
    announce("max size payload", N_MAX)
    A0 = torch.empty(N_MAX, dtype=torch.int32, device="cuda")
    ctx.gen_payload(A0, 131)
    torch.cuda.synchronize()
    host = A0.cpu().numpy()
    yield A0, host
    del A0
    torch.cuda.empty_cache()


def _removals(ctx, k, seed):
    R = torch.empty(k, dtype=torch.int32, device="cuda")
    ctx.gen_uniform_removals(R, N_MAX, seed)
    return R


@pytest.mark.parametrize("k,config", [(1 << 20, None),                               # tuned: mark + fused move/pair
                                      (1 << 14, rz.TileConfig(256, 4, 512, 6, 12, 1)),   # one launch
                                      (1 << 29, None)])  # tuned: dense form, bucketed marks at their 2^29 limit
def test_max_size_redzone(ctx, oracle, big, k, config):
    A0, host = big
    announce("max size redzone", N_MAX, k, config)
    R = _removals(ctx, k, 140 + k.bit_length())
    R_np = R.cpu().numpy()
    A = A0.clone()
    assert ctx.remove(A, R, config=config) == N_MAX - k
    torch.cuda.synchronize()
    got = A.cpu().numpy()
    del A
    want = expected_canonical(oracle, host, R_np)
    m = first_mismatch(got[: N_MAX - k], want[: N_MAX - k])
    assert m < 0, f"first mismatch at {m} (k={k}, cfg={config})"


def test_max_size_partition(ctx, oracle, big):
    A0, host = big
    k = 1 << 19
    announce("max size partition", N_MAX, k)
    R = _removals(ctx, k, 150)
    R_np = R.cpu().numpy()
    A = A0.clone()
    assert ctx.partition(A, R) == N_MAX - k
    torch.cuda.synchronize()
    got = A.cpu().numpy()
    del A
    want = host.copy()
    st, _ = oracle.remove_stable(want, R_np)
    assert st == 0
    m = first_mismatch(got[: N_MAX - k], want[: N_MAX - k])
    assert m < 0, f"first mismatch at {m}"


def test_max_size_gather(ctx, big):
    A0, _ = big
    P = torch.arange(0, N_MAX, 3, dtype=torch.int32, device="cuda")
    p = P.shape[0]
    announce("max size gather", N_MAX, p)
    A = A0.clone()
    assert ctx.gather_preserve(A, P, n=N_MAX) == p
    torch.cuda.synchronize()
    assert torch.equal(A[:p], A0[P.long()])
    assert torch.equal(A[p:], A0[p:]), "the gather wrote past p"


def test_max_size_rec16(ctx, oracle, big):
    """16-byte records at n = 2^31 - 1 (32 GiB, byte offsets past 2^32)."""
    k = 1 << 20
    announce("max size rec16", N_MAX, k)
    A = torch.empty((N_MAX, 4), dtype=torch.int32, device="cuda")
    ctx.gen_payload(A, 161)
    R = _removals(ctx, k, 162)
    torch.cuda.synchronize()
    host = A.cpu().numpy()
    R_np = R.cpu().numpy()
    assert ctx.remove(A, R) == N_MAX - k
    torch.cuda.synchronize()
    got = A.cpu().numpy()
    del A
    torch.cuda.empty_cache()
    want = expected_canonical(oracle, host, R_np)
    del host
    m = first_mismatch(got[: N_MAX - k], want[: N_MAX - k])
    assert m < 0, f"first mismatch at {m}"
