"""The reference's own test shapes and edge values, on every launch form.

- Red-zone-share lists (TestLists_t::Update(n, k, r), DataStruct.hxx:97-122),
  r over the reference's sweep {100, -1, 0, 25, 50, 75} (redzone.hxx:1094),
  at the reference's densities k/n in {2%, 10%, 50%} (redzone.hxx:1114-1121;
  r >= 0 is skipped above 50% there as here).
- Payloads made of the words a marking scheme could take for a mark: the
  reference's GPU sentinel 0x80000000 (redzone.hxx:704), its O(n) marker -1
  (redzone.hxx:594), float NaNs with payloads, -0.0, the infinities, a
  denormal (workloads.EDGE_WORDS), in the head and in the red zone alike.
  Redzone forms are compared bit for bit with the canonical oracle, the O(n)
  partition with the exact stable order, batches with both.
"""
import numpy as np
import pytest
import torch

import redzone_amd as rz
from redzone_amd import workloads as W
from tests.parity import announce, expected_canonical, first_mismatch, gpu_remove_parity

pytestmark = pytest.mark.gpu


def forms(esz):
    """The tuned table, then one shape of every launch form; the dense form's
    1024 x 8 staging fits the 160 KiB of LDS for 4- and 8-byte items only."""
    dw, di = (1024, 8) if esz < 16 else (1024, 4)
    return [None,                                   # the tuned table
            rz.TileConfig(256, 4, 64, 0, 0, 0),     # three kernels
            rz.TileConfig(256, 8, 4096, 3, 0, 0),   # ordered single launch
            rz.TileConfig(256, 8, 512, 4, 0, 0),    # mark + fused move-and-pair
            rz.TileConfig(256, 4, 512, 6, 12, 0),   # one-launch fused
            rz.TileConfig(dw, di, 1024, 5, 11, 2),  # dense, bucketed marks
            rz.TileConfig(dw, di, 1024, 5, 11, 0)]  # dense, atomic marks


@pytest.mark.parametrize("esz", [4, 8, 16])
@pytest.mark.parametrize("frac", [0.02, 0.1, 0.5])
def test_redzone_share_lists(ctx, oracle, esz, frac):
    n = (1 << 20) + 7
    k = int(n * frac)
    A = W.payload(esz, n, seed=40)
    for r in (100, -1, 0, 25, 50, 75):
        R = W.redzone_pct_removals(n, k, r, seed=41 + r)
        want = expected_canonical(oracle, A.numpy(), R.numpy())
        for cfg in forms(esz):
            announce("red zone share", esz, frac, r, cfg)
            gpu_remove_parity(ctx, oracle, A, R, config=cfg, want=want)


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_edge_payloads_redzone(ctx, oracle, esz):
    for n, frac, pattern in [((1 << 20) + 3, 0.01, "uniform"), (1 << 18, 0.3, "tail_heavy"),
                             ((1 << 16) + 1, 0.5, "clustered")]:
        k = int(n * frac)
        A = W.edge_payload(esz, n, seed=50)
        R = W.removals(pattern, n, k, seed=51)
        want = expected_canonical(oracle, A.numpy(), R.numpy())
        for cfg in forms(esz):
            announce("edge payload", esz, n, k, pattern, cfg)
            got, _ = gpu_remove_parity(ctx, oracle, A, R, config=cfg, want=want)
            # the bits, not just the multiset: every sentinel word that was moved kept its value
            assert np.array_equal(got[: n - k].view(np.uint32), want[: n - k].view(np.uint32))


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_edge_payloads_partition(ctx, oracle, esz):
    n = (1 << 20) + 3
    for frac in (0.01, 0.5):
        k = int(n * frac)
        A = W.edge_payload(esz, n, seed=52)
        R = W.uniform_removals(n, k, seed=53)
        announce("edge payload partition", esz, k)
        dA, dR = A.cuda(), R.cuda()
        assert ctx.partition(dA, dR) == n - k
        torch.cuda.synchronize()
        assert ctx.check_errors() == 0
        got = dA.cpu().numpy()
        want = np.delete(A.numpy(), R.numpy().astype(np.int64), axis=0)
        assert np.array_equal(got[: n - k].view(np.uint32), want.view(np.uint32))


def test_edge_payloads_batch(ctx, oracle):
    specs = [(4096 + 17 * i, (4096 + 17 * i) // (3 + i % 5), (4, 8, 16)[i % 3]) for i in range(24)]
    lists, host = [], []
    for i, (n, k, esz) in enumerate(specs):
        A = W.edge_payload(esz, n, seed=60 + i)
        R = W.removals(("uniform", "tail_heavy", "clustered")[i % 3], n, k, seed=90 + i)
        host.append((A.numpy(), R.numpy()))
        lists.append((A.cuda(), R.cuda()))
    announce("edge payload batch")
    ctx.remove_batched(lists)
    torch.cuda.synchronize()
    assert ctx.check_errors() == 0
    for i, ((A0, R0), (dA, _)) in enumerate(zip(host, lists)):
        n, k, _ = specs[i]
        want = expected_canonical(oracle, A0, R0)
        got = dA.cpu().numpy()
        assert first_mismatch(got[: n - k], want[: n - k]) < 0, (i, specs[i])
