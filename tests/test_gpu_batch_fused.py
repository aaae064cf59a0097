"""Batches with fused items (batch variant 4): a list of several slices pairs
its orphans inside its move items, through the list's look-back words and
exchange words, with no PAIR items.  Every list must be array-identical to
what the canonical oracle gives it on its own, on first launches, relaunches,
graph replays and re-planned batches (the MARK items reset the list state
each launch), and lists whose slices are too long for LDS fall back to MOVE
and PAIR items inside the same batch."""
import numpy as np
import pytest
import torch

import redzone_amd as rz
from redzone_amd import workloads as W
from tests.parity import announce
from tests.test_gpu_batched import Lists, _random_specs

pytestmark = pytest.mark.gpu

FUSED = [rz.TileConfig(256, 4, 0, 4, 1, 1), rz.TileConfig(512, 8, 0, 4, 1, 1), rz.TileConfig(512, 8, 0, 4, 2, 2),
         rz.TileConfig(256, 8, 0, 4, 4, 1)]


def test_fused_items_small_lists(ctx, oracle):
    rng = np.random.default_rng(303)
    L = Lists(_random_specs(rng, 300, 60000), seed=8000)
    for cfg in FUSED:
        announce("fused batch small", cfg.as_tuple())
        L.restore()
        b = ctx.batch(L.pairs(), config=cfg)
        b.launch()
        L.check(oracle, f"fused {cfg.as_tuple()}")
        b.close()


def test_fused_items_large_lists_and_fallback(ctx, oracle):
    """Lists of up to 3 M removals run fused items; 9 M removals need slices
    longer than 8,192 pairs and run MOVE and PAIR items in the same batch."""
    specs = [(1 << 22, 1 << 21, 4, "uniform"), ((1 << 22) + 5, 3_000_000, 8, "clustered"),
             (1 << 24, 9_000_000, 4, "uniform"), (300_000, 250_000, 16, "tail_heavy"), (1 << 20, 1, 4, "uniform"),
             (5000, 0, 4, "uniform"), (70_000, 69_999, 8, "uniform")]
    L = Lists(specs, seed=8100)
    for cfg in (rz.TileConfig(512, 8, 0, 4, 1, 1), rz.TileConfig(256, 4, 0, 4, 1, 1)):
        announce("fused batch large", cfg.as_tuple())
        L.restore()
        b = ctx.batch(L.pairs(), config=cfg)
        b.launch()
        L.check(oracle, f"fused large {cfg.as_tuple()}")
        b.close()
    assert ctx.check_errors() == 0


def test_fused_items_config4_shape_relaunch(ctx, oracle):
    """4096 int32 lists, n log-uniform in [2^8, 2^20], k = floor(r n), r ~
    U[0, 0.3]; launched twice (the second launch starts from the first one's
    look-back and exchange words)."""
    rng = np.random.default_rng(4)
    specs = []
    for _ in range(4096):
        n = int(2 ** rng.uniform(8, 20))
        specs.append((n, int(rng.uniform(0, 0.3) * n), 4, "uniform"))
    L = Lists(specs, seed=8200)
    b = ctx.batch(L.pairs(), config=rz.TileConfig(512, 8, 0, 4, 1, 1))
    for rep in range(2):
        announce("fused batch config4", rep)
        L.restore()
        b.launch()
        L.check(oracle, f"fused config4 launch {rep}")
    b.close()
    assert ctx.check_errors() == 0


def test_fused_items_graph_replay(ctx, oracle):
    rng = np.random.default_rng(56)
    L = Lists(_random_specs(rng, 120, 200000), seed=8300)
    b = ctx.batch(L.pairs(), config=rz.TileConfig(512, 8, 0, 4, 1, 1))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        L.restore()
        torch.cuda.synchronize()
        b.capture(stream=stream)
        for rep in range(4):
            announce("fused graph replay", rep)
            L.restore()
            torch.cuda.synchronize()
            b.launch(stream=stream)
            stream.synchronize()
            L.check(oracle, f"fused replay {rep}")
    b.close()


def test_fused_items_replanned_one_shot_and_preserve(ctx, oracle):
    """rz_remove_batched re-plans the context's batch on every call: two
    different list sets in a row (the second one's lists sit where the first
    one's look-back and exchange words were), then RZ_PRESERVE_REMOVALS."""
    old = ctx.batch_config()
    ctx.set_batch_config(rz.TileConfig(256, 4, 0, 4, 1, 1))
    try:
        for seed in (8400, 8500):
            rng = np.random.default_rng(seed)
            L = Lists(_random_specs(rng, 64, 150000), seed=seed)
            announce("fused one shot", seed)
            ctx.remove_batched(L.pairs())
            L.check(oracle, f"fused one-shot {seed}")
        L.restore()
        announce("fused preserve")
        ctx.remove_batched(L.pairs(), flags=rz.RZ_PRESERVE_REMOVALS)
        L.check(oracle, "fused preserve")
        for d, h in zip(L.dev_R, L.host_R):
            assert torch.equal(d.cpu(), h)
    finally:
        ctx.set_batch_config(old)
    assert ctx.check_errors() == 0


def test_fused_items_mixed_stream_sample(ctx, oracle):
    """The batchable lists (fewer than 2^22 removals) of a seeded sample of the
    mixed stream (n log-uniform in [2^12, 2^26], 4/8/16-byte items, three
    removal patterns) in fused batches, each list checked against the oracle."""
    from redzone_amd import mixed as MX
    from tests.parity import expected_canonical, first_mismatch

    specs = [s for s in MX.stream_specs(160, seed=19) if s[1] < MX.BATCH_MAX_K]
    assert {s[2] for s in specs} == {4, 8, 16}
    cfg = rz.TileConfig(512, 8, 0, 4, 1, 1)
    for g in MX.batches(specs, 4 << 30):
        announce("fused mixed sample", g[0], g[-1])
        lists = [MX.build_list(ctx, specs[i], 9100 + 7 * i) for i in g]
        host = [(A.cpu().numpy(), R.cpu().numpy()) for A, R in lists]
        b = ctx.batch(lists, config=cfg)
        b.launch()
        torch.cuda.synchronize()
        b.close()
        assert ctx.check_errors() == 0
        for j, i in enumerate(g):
            n, k = specs[i][0], specs[i][1]
            want = expected_canonical(oracle, host[j][0], host[j][1])
            m = first_mismatch(lists[j][0].cpu().numpy()[: n - k], want[: n - k])
            assert m < 0, (specs[i], m)
        del lists, host
        torch.cuda.empty_cache()


def test_fused_items_orphan_mismatch(ctx, oracle):
    """A list that breaks the contract (its last index listed twice) beside
    valid ones: the fused items see the orphan totals differ and the check
    reports RZ_ERR_ORPHAN_MISMATCH; the next batch on the context is exact."""
    n, k = 1 << 16, 5000
    A = W.unique_u32_payload(n, seed=8600).cuda()
    head = W.uniform_removals(n - k, k - 2, seed=8601)
    R = torch.cat([torch.tensor([n - 1, n - 1], dtype=torch.int32), head]).cuda()
    rng = np.random.default_rng(8602)
    L = Lists(_random_specs(rng, 40, 50000), seed=8603)
    cfg = rz.TileConfig(256, 4, 0, 4, 1, 1)  # 1,024-pair slices: the bad list has 5
    announce("fused mismatch")
    b = ctx.batch(L.pairs() + [(A, R)], config=cfg)
    b.launch()
    b.close()
    with pytest.raises(rz.RedzoneError) as e:
        ctx.check_errors()
    assert e.value.status == rz.RZ_EDEVICE and e.value.error_flags & rz.RZ_ERR_ORPHAN_MISMATCH
    L.restore()
    b = ctx.batch(L.pairs(), config=cfg)
    b.launch()
    L.check(oracle, "fused after mismatch")
    b.close()
    assert ctx.check_errors() == 0
