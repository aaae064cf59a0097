"""Parity of the HIP O(n) stable partition (through the C ABI): the survivors
must equal numpy.delete / the oracle's stable remove item by item."""
import numpy as np
import pytest
import torch

from redzone_amd import workloads as W
from tests.parity import announce

pytestmark = pytest.mark.gpu


def _run(ctx, oracle, A, R):
    n, k = A.shape[0], R.shape[0]
    A_np = A.numpy()
    R_np = R.numpy()
    want = A_np.copy()
    st, _ = oracle.remove_stable(want, R_np, threads=8)
    assert st == 0
    dA, dR = A.cuda(), R.cuda()
    assert ctx.partition(dA, dR) == n - k
    torch.cuda.synchronize()
    got = dA.cpu().numpy()
    assert np.array_equal(got[: n - k], want[: n - k])
    assert torch.equal(dR.cpu(), R), "the partition must not modify the removal list"
    bad, fb = oracle.check(A_np, got, R_np, mode=1)
    assert bad == 0, (bad, fb)


@pytest.mark.parametrize("esz", [4, 8, 16])
@pytest.mark.parametrize("frac", [0.0005, 0.01, 0.3, 0.5, 0.99])
def test_partition_patterns(ctx, oracle, esz, frac):
    n = (1 << 19) + 77
    k = int(n * frac)
    announce("partition", esz, frac)
    _run(ctx, oracle, W.payload(esz, n, seed=3), W.uniform_removals(n, k, seed=4))


def test_partition_ragged(ctx, oracle):
    rng = np.random.default_rng(7)
    for trial in range(60):
        n = int(rng.integers(1, 20000))
        k = int(rng.integers(0, n + 1))
        announce("partition ragged", trial, n, k)
        A = torch.from_numpy(rng.integers(-2 ** 31, 2 ** 31, n, dtype=np.int64).astype(np.int32))
        R = torch.from_numpy(rng.permutation(n)[:k].astype(np.int32))
        _run(ctx, oracle, A, R)


def test_partition_unaligned_list(ctx, oracle):
    """A list that starts one item into a buffer is not 16-byte aligned: the
    scalar tile kernel handles it."""
    n = (1 << 18) + 13
    esz = 4
    full = W.payload(esz, n + 1, seed=5)
    R = W.uniform_removals(n, n // 7, seed=6)
    A_np = full.numpy()[1:].copy()
    want = A_np.copy()
    assert oracle.remove_stable(want, R.numpy(), threads=4)[0] == 0
    dfull = full.cuda()
    view = dfull[1:]
    announce("partition unaligned")
    assert ctx.partition(view, R.cuda()) == n - R.shape[0]
    torch.cuda.synchronize()
    got = dfull.cpu().numpy()[1:]
    assert np.array_equal(got[: n - R.shape[0]], want[: n - R.shape[0]])


def test_partition_repeated_calls(ctx, oracle):
    """Back-to-back calls reuse the bitmap and the look-back state."""
    n = 1 << 20
    for i, frac in enumerate([0.2, 0.001, 0.7, 0.05]):
        announce("partition repeat", i)
        _run(ctx, oracle, W.unique_u32_payload(n, seed=10 + i), W.clustered_removals(n, int(n * frac), seed=20 + i))


def test_partition_full_size(ctx, oracle):
    n = 1 << 28
    k = n // 1000
    announce("partition full 2^28")
    A = W.unique_u32_payload(n, seed=31, device="cuda")
    R = W.uniform_removals(n, k, seed=32, device="cuda")
    A_np = A.cpu().numpy()
    R_np = R.cpu().numpy()
    assert ctx.partition(A, R) == n - k
    torch.cuda.synchronize()
    got = A.cpu().numpy()
    bad, fb = oracle.check(A_np, got, R_np, mode=1)
    assert bad == 0, (bad, fb)


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_partition_bucketed_marks(ctx, oracle, esz):
    """Lists of at least 2^20 items with at least 2^19 removals group their
    marks by bitmap region (bins of 2^15 to 2^19 positions, 256 to 1,024 of
    them from 2^23 items up, several workgroups per bin below); list lengths
    that end inside a region, clustered removals that fill whole regions, the
    bin count changing from call to call, and k = n - 1."""
    for n, k, pattern in [((1 << 23) + 12345, 1 << 20, "uniform"), ((1 << 23) + 12345, 3 << 21, "clustered"),
                          (1 << 20, 1 << 19, "clustered"), ((1 << 21) + 777, 1_500_000, "uniform"),
                          (6_000_001, 4_800_000, "clustered"), ((1 << 22) + 3, 1 << 19, "tail_heavy"),
                          (1 << 23, (1 << 23) - 1, "uniform"), ((3 << 22) + 1, 1 << 22, "tail_heavy")]:
        announce("partition bucketed", esz, n, k, pattern)
        _run(ctx, oracle, W.payload(esz, n, seed=40), W.removals(pattern, n, k, seed=41))


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_partition_tile_counts_change_between_calls(ctx, oracle, esz):
    """Back-to-back calls whose tile counts grow and shrink (partial groups of
    64 tiles at the end, groups that a previous call left in other states):
    the per-tile loaded flags and the per-group words are keyed by call epoch."""
    tile = 1024 * 4 * (16 // esz)
    rng = np.random.default_rng(60 + esz)
    for i, tiles in enumerate([197, 70, 331, 5, 130, 64, 129]):
        n = tiles * tile - int(rng.integers(0, tile // 2))
        for frac in (0.01, 0.5):
            announce("partition tiles", esz, tiles, n, frac)
            _run(ctx, oracle, W.payload(esz, n, seed=61 + i), W.uniform_removals(n, int(n * frac), seed=62 + i))
