"""Parity of the widened entry points around the Redzone path:

- rz_add_remove: append G, then Redzone on A+G (paper p.9 sec. 4.6) equals the
  canonical oracle on the concatenated list;
- rz_gather_preserve: the in-place gather by a preserve list (paper p.14,
  eq. 3) equals numpy's A[P] and, with P the complement of a removal list, the
  oracle's stable O(n) removal (std_parallel_remove.hxx:13-36 semantics).
"""
import numpy as np
import pytest
import torch

import redzone_amd as rz
from redzone_amd import workloads as W
from tests.parity import announce, expected_canonical, first_mismatch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_add_remove(ctx, oracle, esz):
    for n, g, k, pattern, in_g in [(1 << 16, 5000, 3000, "uniform", False),   # g >= k: all moves of added items
                                   (1 << 16, 1000, 6000, "uniform", False),   # g < k: the red zone spans A and G
                                   (1 << 18, 20000, 20000, "clustered", True),  # removals inside G as well
                                   (5000, 0, 700, "tail_heavy", False),       # nothing added: plain removal
                                   (5000, 800, 0, "uniform", False)]:         # nothing removed: plain append
        m = n + g
        A = W.payload(esz, m + 17, seed=50)  # capacity beyond n + g
        G = W.payload(esz, g, seed=51) if g else A[:0].clone()
        R = W.removals(pattern, m if in_g else n, k, seed=52)
        comb = np.concatenate([A.numpy()[:n], G.numpy()])
        want = expected_canonical(oracle, comb, R.numpy())
        announce("add_remove", esz, n, g, k, pattern, in_g)
        dA, dG, dR = A.cuda(), G.cuda(), R.cuda()
        new_len = ctx.add_remove(dA, n, dG, dR)
        torch.cuda.synchronize()
        assert new_len == m - k
        got = dA.cpu().numpy()
        mm = first_mismatch(got[: m - k], want[: m - k])
        assert mm < 0, (n, g, k, mm)


def test_add_remove_rejects_bad_arguments(ctx):
    A = torch.zeros(100, dtype=torch.int32, device="cuda")
    G = torch.zeros(50, dtype=torch.int32, device="cuda")
    R = torch.zeros(0, dtype=torch.int32, device="cuda")
    with pytest.raises(rz.RedzoneError):
        ctx.add_remove(A, 60, G, R)  # capacity 100 < 60 + 50
    with pytest.raises(rz.RedzoneError):
        ctx.add_remove(A, 10, A[10:40], R)  # the added items overlap their slot


def _gather_cases(n, seed):
    rng = np.random.default_rng(seed)
    for frac in (0.01, 0.5, 0.99):
        yield f"uniform {frac}", np.sort(rng.choice(n, int(n * frac), replace=False)).astype(np.int32)
    R = W.clustered_removals(n, n // 3, seed=seed).numpy()
    keep = np.ones(n, bool)
    keep[R] = False
    yield "clustered complement", np.nonzero(keep)[0].astype(np.int32)
    yield "identity", np.arange(n, dtype=np.int32)
    yield "last one", np.array([n - 1], dtype=np.int32)


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_gather_preserve(ctx, esz):
    for n in ((1 << 20) + 3, 3000):
        A = W.payload(esz, n, seed=53)
        A_np = A.numpy()
        for name, P in _gather_cases(n, seed=54):
            announce("gather", esz, n, name, len(P))
            dA = A.cuda()
            assert ctx.gather_preserve(dA, torch.from_numpy(P).cuda(), n=n) == len(P)
            torch.cuda.synchronize()
            got = dA.cpu().numpy()
            want = A_np[P]
            mm = first_mismatch(got[: len(P)], want)
            assert mm < 0, (name, mm)
            # the items past p are untouched or moved-from copies, never out of the list
            assert got.shape == A_np.shape


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_gather_preserve_one_pass(ctx, esz):
    """From p = 2^22 on the gather is one in-place pass ordered by tile flags;
    calls of different p (so different tile counts) follow each other, with a
    small two-pass call between them."""
    n = (1 << 23) + 5
    A = W.payload(esz, n, seed=58)
    A_np = A.numpy()
    rng = np.random.default_rng(59)
    R = W.clustered_removals(n, n // 3, seed=60).numpy()
    keep = np.ones(n, bool)
    keep[R] = False
    cases = [("uniform 0.99", np.sort(rng.choice(n, int(n * 0.99), replace=False)).astype(np.int32)),
             ("uniform 0.5", np.sort(rng.choice(n, n // 2, replace=False)).astype(np.int32)),
             ("small, two passes", np.sort(rng.choice(n, 5000, replace=False)).astype(np.int32)),
             ("clustered complement", np.nonzero(keep)[0].astype(np.int32)),
             ("identity", np.arange(n, dtype=np.int32))]
    for name, P in cases:
        announce("gather one pass", esz, n, name, len(P))
        dA = A.cuda()
        assert ctx.gather_preserve(dA, torch.from_numpy(P).cuda(), n=n) == len(P)
        torch.cuda.synchronize()
        mm = first_mismatch(dA.cpu().numpy()[: len(P)], A_np[P])
        assert mm < 0, (name, mm)


def test_gather_preserve_equals_stable_removal(ctx, oracle):
    """With P the ascending complement of R, the gather is the stable O(n)
    removal: the oracle's restatement of std::remove on the marked copy."""
    n = 1 << 19
    for pattern, frac in (("uniform", 0.1), ("tail_heavy", 0.5), ("clustered", 0.01)):
        A = W.payload(4, n, seed=55)
        R = W.removals(pattern, n, int(n * frac), seed=56)
        keep = np.ones(n, bool)
        keep[R.numpy()] = False
        P = np.nonzero(keep)[0].astype(np.int32)
        want = A.numpy().copy()
        st, _ = oracle.remove_stable(want, R.numpy())
        assert st == 0
        announce("gather vs stable removal", pattern, frac)
        dA = A.cuda()
        ctx.gather_preserve(dA, torch.from_numpy(P).cuda(), flags=rz.RZ_VALIDATE)
        torch.cuda.synchronize()
        assert first_mismatch(dA.cpu().numpy()[: len(P)], want[: len(P)]) < 0


def test_gather_preserve_validation(ctx):
    n = 1000
    A = W.payload(4, n, seed=57).cuda()
    A0 = A.clone()
    for bad, status in (([5, 3, 9], 1), ([5, 5, 9], 3), ([5, 9, n], 2)):
        announce("gather validate", bad)
        with pytest.raises(rz.RedzoneError) as e:
            ctx.gather_preserve(A, torch.tensor(bad, dtype=torch.int32).cuda(), flags=rz.RZ_VALIDATE)
        assert e.value.status == status
    torch.cuda.synchronize()
    assert torch.equal(A, A0), "rejected calls must not touch the data"
