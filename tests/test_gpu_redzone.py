"""Parity of the HIP Redzone path (through the C ABI) with the CPU oracle.

Bar: the survivors d[0, n-k) are array-identical to the canonical sequential
oracle (which the oracle tests pin against the paper's worked examples and the
Redzone_CPU restatement), and satisfy every invariant: length n-k, survivor
multiset with payload bits, in-place invariant, no removed item left.
"""
import json
import os

import numpy as np
import pytest
import torch

import redzone_amd as rz
from redzone_amd import workloads as W
from tests.parity import announce, expected_canonical, first_mismatch, gpu_remove_parity

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


def _payload(esz, n, seed):
    return W.payload(esz, n, seed)


def test_known_answers(ctx, oracle):
    for case in json.load(open(GOLDEN))["cases"]:
        announce("known_answer", case["name"])
        n = case["n"]
        if n == 0:
            continue
        A = torch.arange(n, dtype=torch.int32)
        R = torch.tensor(case["R"], dtype=torch.int32)
        got, _ = gpu_remove_parity(ctx, oracle, A, R)
        assert got[: n - len(case["R"])].tolist() == case["expect"]


def test_config1_index_payload(ctx, oracle):
    """n=2^20 int32 A[i]=i, k=2^14 uniform distinct removals, unsorted; also
    checked against the Redzone_CPU restatement."""
    n, k = 1 << 20, 1 << 14
    announce("config1", n, k)
    A = W.index_payload(n)
    R = W.uniform_removals(n, k, seed=1)
    got, _ = gpu_remove_parity(ctx, oracle, A, R)
    A2 = A.numpy().copy()
    st, _ = oracle.redzone_cpu(A2, R.numpy().copy(), threads=7)
    assert st == 0
    assert np.array_equal(got[: n - k], A2[: n - k])


@pytest.mark.parametrize("esz", [4, 8, 16])
@pytest.mark.parametrize("pattern", ["uniform", "clustered", "tail_heavy"])
@pytest.mark.parametrize("frac", [1e-3, 1e-2, 0.1, 0.5, 0.9])
def test_patterns_and_sizes(ctx, oracle, esz, pattern, frac):
    n = 1 << 18
    k = int(n * frac)
    announce("pattern", esz, pattern, frac)
    A = _payload(esz, n, seed=11)
    R = W.removals(pattern, n, k, seed=12)
    gpu_remove_parity(ctx, oracle, A, R)


CONFIGS = [rz.TileConfig(wg, it, ms, 0, 0, 0) for wg in (256, 512) for it in (2, 4, 8, 16) for ms in (1, 3, 64, 4096)]


@pytest.mark.parametrize("esz", [4, 16])
def test_launch_shape_invariance(ctx, oracle, esz):
    """Every launch shape gives the identical array (canonical pairing is
    independent of the slicing)."""
    n = (1 << 17) + 123
    k = int(n * 0.3)
    A = _payload(esz, n, seed=21)
    R = W.uniform_removals(n, k, seed=22)
    want = expected_canonical(oracle, A.numpy(), R.numpy())
    for cfg in CONFIGS:
        announce("shape", esz, cfg)
        gpu_remove_parity(ctx, oracle, A, R, config=cfg, want=want)


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_cooperative_variant(ctx, oracle, esz):
    """The single cooperative launch (in-kernel grid barriers) gives the same array."""
    for n, frac, ms in [((1 << 18) + 5, 0.01, 4096), ((1 << 18) + 5, 0.4, 64), (1 << 16, 0.9, 7), (1 << 20, 0.05, 4096)]:
        k = int(n * frac)
        A = _payload(esz, n, seed=31)
        R = W.tail_heavy_removals(n, k, seed=32)
        for wg, it in ((256, 4), (512, 8)):
            cfg = rz.TileConfig(wg, it, ms, 1, 0, 0)
            announce("coop", esz, n, k, cfg)
            gpu_remove_parity(ctx, oracle, A, R, config=cfg)


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_ordered_variant(ctx, oracle, esz):
    """The ordered single launch (roles by workgroup id, in-kernel counters) gives the same array."""
    for n, frac, ms, pattern in [((1 << 18) + 5, 0.01, 4096, "uniform"), ((1 << 18) + 5, 0.4, 64, "tail_heavy"),
                                 (1 << 16, 0.9, 7, "clustered"), (1 << 20, 0.05, 4096, "tail_heavy"),
                                 (3000, 0.5, 4096, "uniform"), (70000, 0.001, 1, "uniform")]:
        k = int(n * frac)
        A = _payload(esz, n, seed=33)
        R = W.removals(pattern, n, k, seed=34)
        want = expected_canonical(oracle, A.numpy(), R.numpy())
        for wg, it in ((256, 4), (256, 16), (512, 8)):
            cfg = rz.TileConfig(wg, it, ms, 3, 0, 0)
            announce("ordered", esz, n, k, pattern, cfg)
            gpu_remove_parity(ctx, oracle, A, R, config=cfg, want=want)


@pytest.mark.parametrize("variant", [4, 6])
@pytest.mark.parametrize("esz", [4, 8, 16])
def test_fused_variant(ctx, oracle, esz, variant):
    """Mark, then one move kernel that pairs orphans through look-back and
    exchange words (variant 4), or both in one launch with the mark workgroups
    first (variant 6): the same array as the three-kernel form.  Consecutive
    calls alternate the look-back status halves, and the slice count changes
    between calls, so a half left dirty (or a mark counter not reset) would
    show up here."""
    cases = [((1 << 18) + 5, 0.01, 4096, "uniform"), ((1 << 18) + 5, 0.4, 64, "tail_heavy"),
             (1 << 16, 0.9, 7, "clustered"), (1 << 20, 0.05, 4096, "tail_heavy"), (3000, 0.5, 4096, "uniform"),
             (70000, 0.001, 1, "uniform"), ((1 << 20) + 77, 0.3, 4096, "uniform"), (1 << 21, 0.5, 512, "clustered")]
    for n, frac, ms, pattern in cases:
        k = int(n * frac)
        A = _payload(esz, n, seed=36)
        R = W.removals(pattern, n, k, seed=37)
        want = expected_canonical(oracle, A.numpy(), R.numpy())
        for wg, it in ((256, 2), (256, 4), (256, 16), (512, 8)):
            cfg = rz.TileConfig(wg, it, ms, variant, 0, 0)
            announce("fused", esz, n, k, pattern, cfg)
            gpu_remove_parity(ctx, oracle, A, R, config=cfg, want=want)


@pytest.mark.parametrize("variant", [4, 6])
def test_fused_variant_flags(ctx, oracle, variant):
    """RZ_STATS, RZ_PRESERVE_REMOVALS and RZ_SORTED on the fused forms; and a
    slice too long for LDS (runs as the three-kernel form)."""
    n = 1 << 20
    k = n // 10
    A = W.index_payload(n)
    R = W.tail_heavy_removals(n, k, seed=38)
    cfg = rz.TileConfig(256, 4, 4096, variant, 0, 0)
    announce("fused stats")
    _, dR = gpu_remove_parity(ctx, oracle, A, R, config=cfg, flags=rz.RZ_STATS | rz.RZ_PRESERVE_REMOVALS)
    s = ctx.last_stats()
    Rn = R.numpy()
    sz = n - k
    dead = np.zeros(k, bool)
    dead[Rn[Rn >= sz] - sz] = True
    assert s.tail_removals == int((Rn >= sz).sum())
    assert s.orphan_pairs == int(((Rn < sz) & dead).sum())
    assert s.error_flags == 0
    assert torch.equal(dR.cpu(), R)
    announce("fused sorted")
    Rs = torch.sort(R).values.contiguous()
    gpu_remove_parity(ctx, oracle, A, Rs, config=cfg, flags=rz.RZ_SORTED)
    announce("fused long slices")
    gpu_remove_parity(ctx, oracle, A, R, config=rz.TileConfig(256, 4, 1, variant, 0, 0))


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_dense_variant(ctx, oracle, esz):
    """Region-bucketed form: head writes grouped by destination bin, dest orphans
    carrying their rank, then a per-bin scatter: the same array again."""
    cases = [((1 << 18) + 5, 0.01, 4096, 9, "uniform"), ((1 << 18) + 5, 0.4, 64, 5, "tail_heavy"),
             (1 << 16, 0.9, 7, 10, "clustered"), (1 << 20, 0.25, 4096, 9, "uniform"), (3000, 0.5, 4096, 9, "uniform"),
             (70000, 0.001, 1, 6, "uniform"), ((1 << 21) + 77, 0.5, 512, 10, "clustered"),
             ((1 << 20) + 3, 0.1, 1024, 8, "tail_heavy")]
    for n, frac, ms, lgb, pattern in cases:
        k = int(n * frac)
        A = _payload(esz, n, seed=39)
        R = W.removals(pattern, n, k, seed=40)
        want = expected_canonical(oracle, A.numpy(), R.numpy())
        shapes = [(256, 4, 0), (512, 8, 2), (1024, 4, 2)]
        if esz < 16:  # the tuned table's 1024 x 8 rows (LDS at its 160 KiB limit for 8-byte items)
            shapes += [(1024, 8, 0), (1024, 8, 2)]
        for wg, it, marks in shapes:
            cfg = rz.TileConfig(wg, it, ms, 5, lgb, marks)
            announce("dense", esz, n, k, pattern, cfg)
            gpu_remove_parity(ctx, oracle, A, R, config=cfg, want=want)


def test_dense_variant_flags(ctx, oracle):
    n = 1 << 20
    k = n // 4
    A = W.index_payload(n)
    R = W.uniform_removals(n, k, seed=41)
    cfg = rz.TileConfig(512, 8, 1024, 5, 9, 2)
    announce("dense stats")
    _, dR = gpu_remove_parity(ctx, oracle, A, R, config=cfg, flags=rz.RZ_STATS | rz.RZ_PRESERVE_REMOVALS)
    s = ctx.last_stats()
    Rn = R.numpy()
    sz = n - k
    dead = np.zeros(k, bool)
    dead[Rn[Rn >= sz] - sz] = True
    assert s.tail_removals == int((Rn >= sz).sum())
    assert s.orphan_pairs == int(((Rn < sz) & dead).sum())
    assert s.error_flags == 0
    assert torch.equal(dR.cpu(), R)
    announce("dense sorted")
    gpu_remove_parity(ctx, oracle, A, torch.sort(R).values.contiguous(), config=cfg, flags=rz.RZ_SORTED)


def test_ordered_variant_stats_and_preserve(ctx, oracle):
    n = 1 << 18
    k = n // 10
    A = W.index_payload(n)
    R = W.tail_heavy_removals(n, k, seed=35)
    cfg = rz.TileConfig(256, 8, 4096, 3, 0, 0)
    announce("ordered stats")
    _, dR = gpu_remove_parity(ctx, oracle, A, R, config=cfg, flags=rz.RZ_STATS | rz.RZ_PRESERVE_REMOVALS)
    s = ctx.last_stats()
    Rn = R.numpy()
    assert s.tail_removals == int((Rn >= n - k).sum())
    assert s.error_flags == 0
    assert torch.equal(dR.cpu(), R)


def test_ragged_small(ctx, oracle):
    rng = np.random.default_rng(41)
    for trial in range(120):
        n = int(rng.integers(1, 700))
        k = int(rng.integers(0, n + 1))
        esz = int(rng.choice([4, 8, 16]))
        announce("ragged", trial, n, k, esz)
        A = torch.from_numpy(rng.integers(-2 ** 31, 2 ** 31, size=(n, esz // 4), dtype=np.int64).astype(np.int32))
        A = A.reshape(n) if esz == 4 else A
        R = torch.from_numpy(rng.permutation(n)[:k].astype(np.int32))
        cfg = rz.TileConfig(int(rng.choice([256, 512])), int(rng.choice([2, 4, 8, 16])), int(rng.integers(1, 50)), 0)
        gpu_remove_parity(ctx, oracle, A, R, config=cfg if trial % 2 else None)


@pytest.mark.parametrize("case", ["k1", "k_n_minus_1", "tail_only", "head_only", "k0", "kn"])
def test_edges(ctx, oracle, case):
    n = 100003
    announce("edge", case)
    A = W.unique_u32_payload(n, seed=51)
    if case == "k1":
        R = torch.tensor([777], dtype=torch.int32)
    elif case == "k_n_minus_1":
        R = W.uniform_removals(n, n - 1, seed=52)
    elif case == "tail_only":
        R = torch.flip(torch.arange(n - 5000, n, dtype=torch.int32), [0])
    elif case == "head_only":
        R = W.uniform_removals(n // 2, 5000, seed=53)
    elif case == "k0":
        R = torch.empty(0, dtype=torch.int32)
    else:
        R = W.uniform_removals(n, n, seed=54)
    gpu_remove_parity(ctx, oracle, A, R)


def test_preserve_removals(ctx, oracle):
    n = 1 << 19
    k = n // 20
    announce("preserve", n, k)
    A = W.unique_u32_payload(n, seed=61)
    R = W.uniform_removals(n, k, seed=62)
    _, dR = gpu_remove_parity(ctx, oracle, A, R, flags=rz.RZ_PRESERVE_REMOVALS)
    assert torch.equal(dR.cpu(), R)


def test_validate_rejects_bad_lists(ctx, oracle):
    n = 5000
    A = W.unique_u32_payload(n, seed=71)
    dA = A.cuda()
    announce("validate dup")
    with pytest.raises(rz.RedzoneError) as e:
        ctx.remove(dA, torch.tensor([5, 9, 5], dtype=torch.int32).cuda(), flags=rz.RZ_VALIDATE)
    assert e.value.status == 3
    announce("validate range")
    with pytest.raises(rz.RedzoneError) as e:
        ctx.remove(dA, torch.tensor([5, n], dtype=torch.int32).cuda(), flags=rz.RZ_VALIDATE)
    assert e.value.status == 2
    with pytest.raises(rz.RedzoneError):
        ctx.remove(dA, torch.tensor([-1], dtype=torch.int32).cuda(), flags=rz.RZ_VALIDATE)
    torch.cuda.synchronize()
    assert torch.equal(dA.cpu(), A), "rejected calls must not touch the data"
    announce("validate ok")
    R = W.uniform_removals(n, 300, seed=72)
    gpu_remove_parity(ctx, oracle, A, R, flags=rz.RZ_VALIDATE)


@pytest.mark.parametrize("n,k,pattern", [(1 << 20, 10_000, "uniform"), (1 << 20, 4096 * 3, "tail_heavy"),
                                         (1 << 18, (1 << 18) // 2 + 7, "clustered"), (70_001, 1, "uniform")])
def test_sorted_flag(ctx, oracle, n, k, pattern):
    """RZ_SORTED reads only the red-zone suffix of an ascending list in the mark
    step; the survivors must be the same as without the flag."""
    announce("sorted", n, k, pattern)
    A = W.unique_u32_payload(n, seed=91)
    R = getattr(W, f"{pattern}_removals")(n, k, seed=92)
    R = torch.sort(R).values.contiguous()
    gpu_remove_parity(ctx, oracle, A, R, flags=rz.RZ_SORTED)
    gpu_remove_parity(ctx, oracle, A, R, flags=rz.RZ_SORTED | rz.RZ_VALIDATE)


def test_sorted_flag_validate_rejects_unsorted(ctx):
    n = 5000
    A = W.unique_u32_payload(n, seed=93)
    dA = A.cuda()
    announce("sorted validate")
    with pytest.raises(rz.RedzoneError) as e:
        ctx.remove(dA, torch.tensor([5, 900, 17], dtype=torch.int32).cuda(), flags=rz.RZ_SORTED | rz.RZ_VALIDATE)
    assert e.value.status == 1
    torch.cuda.synchronize()
    assert torch.equal(dA.cpu(), A)


def test_stats(ctx, oracle):
    n = 1 << 18
    k = n // 10
    announce("stats", n, k)
    A = W.index_payload(n)
    R = W.tail_heavy_removals(n, k, seed=81)
    gpu_remove_parity(ctx, oracle, A, R, flags=rz.RZ_STATS)
    s = ctx.last_stats()
    Rn = R.numpy()
    sz = n - k
    dead = np.zeros(k, bool)
    dead[Rn[Rn >= sz] - sz] = True
    dst_orphans = int(((Rn < sz) & dead).sum())
    assert s.tail_removals == int((Rn >= sz).sum())
    assert s.orphan_pairs == dst_orphans
    assert s.error_flags == 0


def test_full_size_u32_one_percent(ctx, oracle):
    """n=2^28 unique uint32 payloads, k/n=1% uniform: array-exact against the oracle."""
    n = 1 << 28
    k = n // 100
    announce("full 2^28 1%")
    A = W.unique_u32_payload(n, seed=91, device="cuda")
    R = W.uniform_removals(n, k, seed=92, device="cuda")
    A_np = A.cpu().numpy()
    R_np = R.cpu().numpy()
    ctx.remove(A, R)
    torch.cuda.synchronize()
    got = A.cpu().numpy()
    want = expected_canonical(oracle, A_np, R_np)
    assert np.array_equal(got[: n - k], want[: n - k])
    bad, fb = oracle.check(A_np, got, R_np)
    assert bad == 0, (bad, fb)


@pytest.mark.parametrize("order", ["sorted", "permuted"])
def test_full_size_rec16(ctx, oracle, order):
    """n=2^27 16-byte records, k=n/100 uniform, removal list sorted / permuted:
    array-exact, float fields compared as bits."""
    n = 1 << 27
    k = n // 100
    announce("full rec16", order)
    A = W.rec16_payload(n, seed=101, device="cuda")
    R = W.uniform_removals(n, k, seed=102, device="cuda")
    if order == "sorted":
        R = torch.sort(R).values
    A_np = A.cpu().numpy()
    R_np = R.cpu().numpy()
    ctx.remove(A, R)
    torch.cuda.synchronize()
    got = A.cpu().numpy()
    want = expected_canonical(oracle, A_np, R_np)
    assert np.array_equal(got[: n - k], want[: n - k])


def _torch_tail_bits(R, n, k):
    """The k-bit tail bitmap built with torch ops (distinct removals: summing the
    bit values of a word is the same as or-ing them)."""
    sz = n - k
    z = (R[R >= sz] - sz).long()
    words = torch.zeros((k + 31) // 32, dtype=torch.int64, device=R.device)
    words.index_add_(0, z >> 5, torch.ones_like(z) << (z & 31))
    return ((words + (1 << 31)) % (1 << 32) - (1 << 31)).to(torch.int32)


@pytest.mark.parametrize("esz", [4, 16])
def test_premarked(ctx, oracle, esz):
    """Producer-side marks: rz_mark_removals equals the bitmap built with torch,
    rz_remove_premarked gives the canonical array for every launch form (the
    single-launch forms run as the fused form) and leaves the bitmap zero."""
    for n, frac, pattern in [((1 << 18) + 5, 0.01, "uniform"), (1 << 18, 0.4, "tail_heavy"), (70000, 0.001, "uniform"),
                             (1 << 20, 0.3, "clustered")]:
        k = int(n * frac)
        A = _payload(esz, n, seed=42)
        R = W.removals(pattern, n, k, seed=43)
        want = expected_canonical(oracle, A.numpy(), R.numpy())
        dR = R.cuda()
        ref_bits = _torch_tail_bits(dR, n, k)
        bits = torch.zeros((k + 31) // 32, dtype=torch.int32, device="cuda")
        announce("mark_removals", esz, n, k, pattern)
        ctx.mark_removals(dR, n, bits)
        torch.cuda.synchronize()
        assert torch.equal(bits, ref_bits)
        for cfg in (None, rz.TileConfig(256, 4, 64, 0, 0, 0), rz.TileConfig(256, 4, 4096, 3, 0, 0),
                    rz.TileConfig(256, 8, 4096, 4, 0, 0), rz.TileConfig(256, 8, 4096, 6, 0, 0),
                    rz.TileConfig(512, 8, 1024, 5, 9, 0),
                    rz.TileConfig(512, 4, 1024, 5, 9, 2)):
            announce("premarked", esz, n, k, pattern, cfg)
            bits.copy_(ref_bits)
            dA = A.cuda()
            dR = R.cuda()
            assert ctx.remove_premarked(dA, dR, bits, config=cfg) == n - k
            torch.cuda.synchronize()
            got = dA.cpu().numpy()
            m = first_mismatch(got[: n - k], want[: n - k])
            assert m < 0, (cfg, m)
            assert int(bits.abs().sum().item()) == 0, "the bitmap must be left zero"
    announce("premarked k == n")
    n = 4096
    A = _payload(4, n, seed=44).cuda()
    R = torch.randperm(n, dtype=torch.int64).to(torch.int32).cuda()
    bits = _torch_tail_bits(R, n, n)
    assert ctx.remove_premarked(A, R, bits) == 0
    torch.cuda.synchronize()
    assert int(bits.abs().sum().item()) == 0


@pytest.mark.parametrize("variant", [4, 6])
@pytest.mark.parametrize("esz", [4, 16])
def test_any_pairing(ctx, oracle, esz, variant):
    """RZ_ANY_PAIRING (early pairing, paper p.14): a slice pairs its own
    orphans first and only the rest go to the global exchange.  The array may
    differ from the canonical one, so the check is the Redzone contract
    (length, survivors with their payload bits, every kept item outside the
    tail in place), and the exchange sees fewer pairs than the canonical
    pairing has."""
    for n, frac, ms, pattern in [((1 << 18) + 5, 0.3, 64, "uniform"), (1 << 20, 0.05, 512, "tail_heavy"),
                                 (70000, 0.5, 7, "clustered"), (1 << 16, 0.9, 4096, "uniform"),
                                 (3000, 0.4, 1, "uniform")]:
        k = int(n * frac)
        A = _payload(esz, n, seed=70)
        R = W.removals(pattern, n, k, seed=71)
        Rn = R.numpy()
        sz = n - k
        dead = np.zeros(k, bool)
        dead[Rn[Rn >= sz] - sz] = True
        canonical_pairs = int(((Rn < sz) & dead).sum())
        announce("any pairing", esz, variant, n, k, pattern)
        dA, dR = A.cuda(), R.cuda()
        cfg = rz.TileConfig(256, 4, ms, variant, 0, 0)
        assert ctx.remove(dA, dR, config=cfg, flags=rz.RZ_ANY_PAIRING | rz.RZ_STATS) == n - k
        s = ctx.last_stats()
        got = dA.cpu().numpy()
        bad, fb = oracle.check(A.numpy(), got, Rn, mode=0)
        assert bad == 0, (n, k, pattern, bad, fb)
        assert s.error_flags == 0
        assert s.orphan_pairs <= canonical_pairs
        if ms > 1 and canonical_pairs > 1000:
            assert s.orphan_pairs < canonical_pairs  # slices with both kinds paired some themselves


@pytest.mark.parametrize("variant", [4, 6])
def test_stage_stamps(ctx, oracle, variant):
    """RZ_PROFILE on the fused forms: every workgroup stamps its stage
    boundaries; the timeline has the launch's workgroup counts and is ordered
    stage by stage; a call of another form leaves no stamps to read."""
    n = 1 << 22
    k = n // 100
    A = W.unique_u32_payload(n, seed=95)
    R = W.uniform_removals(n, k, seed=96)
    cfg = rz.TileConfig(256, 4, 512, variant, 12 if variant == 6 else 0, 0)
    announce("stage stamps", variant)
    gpu_remove_parity(ctx, oracle, A, R, config=cfg, flags=rz.RZ_PROFILE)
    st = ctx.last_stages()
    assert st["slices"] == (k + 1023) // 1024  # one step of 256 x 4 pairs per slice at this k
    assert st["mark_groups"] == ((k + 4095) // 4096 if variant == 6 else 0)
    assert 0 <= st["first_slice_start"] <= st["last_slice_start"]
    assert 0 < st["lookback_end"] <= st["moves_end"] <= st["exchange_end"] <= st["exit"] < 10_000
    if variant == 6:
        assert 0 < st["marks_done"] <= st["ready_raised"] <= st["ready_wait_end"] <= st["lookback_end"]
    gpu_remove_parity(ctx, oracle, A, R, config=rz.TileConfig(256, 4, 64, 0, 0, 0), flags=rz.RZ_PROFILE)
    with pytest.raises(rz.RedzoneError):
        ctx.last_stages()
