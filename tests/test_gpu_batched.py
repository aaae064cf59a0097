"""Parity of the persistent work-queue kernel (batched lists, graph replay,
and the single-list queue variant) with the CPU oracle: every list must be
array-identical to what the canonical oracle gives it on its own."""
import numpy as np
import pytest
import torch

import redzone_amd as rz
from redzone_amd import workloads as W
from tests.parity import announce, expected_canonical, first_mismatch, gpu_remove_parity

pytestmark = pytest.mark.gpu


class Lists:
    """Many lists packed into one device buffer per element size."""

    def __init__(self, specs, seed):
        # specs: list of (n, k, esz, pattern)
        self.specs = specs
        self.host_A, self.host_R = [], []
        for i, (n, k, esz, pattern) in enumerate(specs):
            self.host_A.append(W.payload(esz, n, seed=seed + 3 * i))
            self.host_R.append(W.removals(pattern, n, k, seed=seed + 3 * i + 1))
        self.dev_A = [a.cuda() for a in self.host_A]
        self.dev_R = [r.cuda() for r in self.host_R]
        self.want = [expected_canonical_np(a, r) for a, r in zip(self.host_A, self.host_R)]

    def restore(self):
        for d, h in zip(self.dev_A, self.host_A):
            d.copy_(h.cuda(non_blocking=False))
        for d, h in zip(self.dev_R, self.host_R):
            d.copy_(h.cuda(non_blocking=False))

    def pairs(self):
        return list(zip(self.dev_A, self.dev_R))

    def check(self, oracle, label):
        torch.cuda.synchronize()
        for i, (n, k, esz, pattern) in enumerate(self.specs):
            got = self.dev_A[i].cpu().numpy()
            m = first_mismatch(got[: n - k], self.want[i][: n - k])
            assert m < 0, f"{label}: list {i} (n={n}, k={k}, esz={esz}, {pattern}) differs at {m}"


def expected_canonical_np(A, R):
    from oracle import pyoracle

    return expected_canonical(pyoracle, A.numpy(), R.numpy())


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_queue_variant_single_list(ctx, oracle, esz):
    for n, frac, pattern in [((1 << 18) + 9, 0.01, "uniform"), ((1 << 18) + 9, 0.3, "tail_heavy"),
                             (1 << 16, 0.9, "clustered"), (5000, 0.2, "uniform"), (1 << 20, 0.05, "uniform")]:
        k = int(n * frac)
        A = W.payload(esz, n, seed=5)
        R = W.removals(pattern, n, k, seed=6)
        want = expected_canonical(oracle, A.numpy(), R.numpy())
        for wg, it, steps in [(256, 4, 1), (256, 8, 4), (512, 4, 2), (512, 8, 16)]:
            cfg = rz.TileConfig(wg, it, 0, 2, steps, 1)
            announce("queue single", esz, n, k, pattern, cfg)
            gpu_remove_parity(ctx, oracle, A, R, config=cfg, want=want)


def test_queue_variant_repeated(ctx, oracle):
    """Back-to-back single-list queue launches (the queue resets itself)."""
    n = 1 << 19
    cfg = rz.TileConfig(256, 8, 0, 2, 2, 1)
    for i, frac in enumerate([0.2, 0.0001, 0.6, 0.01, 0.01]):
        announce("queue repeat", i)
        gpu_remove_parity(ctx, oracle, W.unique_u32_payload(n, seed=i), W.uniform_removals(n, int(n * frac), seed=9 + i),
                          config=cfg)


def _random_specs(rng, count, nmax, esz_choices=(4, 8, 16)):
    specs = []
    for i in range(count):
        n = int(rng.integers(1, nmax))
        r = rng.random()
        k = 0 if r < 0.05 else n if r < 0.08 else int(rng.integers(0, n + 1))
        specs.append((n, k, int(rng.choice(esz_choices)), ["uniform", "clustered", "tail_heavy"][i % 3]))
    return specs


def test_batch_mixed_small(ctx, oracle):
    rng = np.random.default_rng(101)
    specs = _random_specs(rng, 300, 60000)
    L = Lists(specs, seed=1000)
    for cfg in [None, rz.TileConfig(256, 4, 0, 2, 1, 1), rz.TileConfig(512, 8, 0, 2, 8, 3)]:
        announce("batch mixed", cfg)
        L.restore()
        b = ctx.batch(L.pairs(), config=cfg)
        b.launch()
        L.check(oracle, f"batch {cfg}")
        b.close()


def test_batch_config4_shape(ctx, oracle):
    """4096 int32 lists, n log-uniform in [2^8, 2^20], k = floor(r n), r ~ U[0, 0.3]."""
    rng = np.random.default_rng(4)
    specs = []
    for i in range(4096):
        n = int(2 ** rng.uniform(8, 20))
        k = int(rng.uniform(0, 0.3) * n)
        specs.append((n, k, 4, "uniform"))
    announce("batch config4 setup")
    L = Lists(specs, seed=7000)
    announce("batch config4 launch")
    b = ctx.batch(L.pairs())
    info = b.info()
    b.launch()
    L.check(oracle, "config4")
    announce("batch config4 relaunch", info)
    L.restore()
    b.launch()
    L.check(oracle, "config4 relaunch")
    b.close()


def test_batch_fewer_lists_than_queues(ctx, oracle):
    """One queue per XCD: with fewer multi-slice lists than queues most
    workgroups find their own queue empty and take another queue's items."""
    for specs in ([(1 << 20, 1 << 17, 4, "uniform")],
                  [(1 << 20, 1 << 17, 4, "uniform"), (1 << 19, 5000, 8, "clustered"),
                   (300000, 250000, 16, "tail_heavy"), (5000, 0, 4, "uniform")]):
        L = Lists(specs, seed=4000 + len(specs))
        for cfg in [None, rz.TileConfig(256, 4, 0, 2, 1, 1)]:
            announce("batch few lists", len(specs), cfg)
            L.restore()
            b = ctx.batch(L.pairs(), config=cfg)
            b.launch()
            L.check(oracle, f"few lists {len(specs)} {cfg}")
            b.close()


def test_batch_graph_replay(ctx, oracle):
    """Capture once, replay several times with eager restores and syncs in between;
    every replay must equal the oracle."""
    rng = np.random.default_rng(55)
    specs = _random_specs(rng, 120, 200000)
    L = Lists(specs, seed=2000)
    b = ctx.batch(L.pairs())
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        L.restore()
        torch.cuda.synchronize()
        b.capture(stream=stream)
        for rep in range(4):
            announce("graph replay", rep)
            L.restore()
            torch.cuda.synchronize()
            b.launch(stream=stream)
            stream.synchronize()
            L.check(oracle, f"replay {rep}")
    b.close()


def test_remove_batched_one_shot_and_preserve(ctx, oracle):
    rng = np.random.default_rng(77)
    specs = _random_specs(rng, 64, 100000)
    L = Lists(specs, seed=3000)
    announce("one shot")
    ctx.remove_batched(L.pairs())
    L.check(oracle, "one-shot")
    announce("preserve")
    L.restore()
    b = ctx.batch(L.pairs(), flags=rz.RZ_PRESERVE_REMOVALS)
    b.launch()
    L.check(oracle, "preserve")
    for d, h in zip(L.dev_R, L.host_R):
        assert torch.equal(d.cpu(), h)
    b.close()


def test_batch_spread_queues(ctx, oracle):
    """Batch config max_slices 16: the items of every list of several slices
    are dealt over all eight queues (every XCD on the same list), with
    MOVE/PAIR items, fused items and a one-queue config beside it; small lists,
    large lists and a relaunch, each list against the oracle."""
    rng = np.random.default_rng(1616)
    specs = _random_specs(rng, 200, 80000)
    specs += [(1 << 22, 1 << 21, 4, "uniform"), ((1 << 21) + 7, 900_000, 16, "clustered"),
              (1 << 20, 300_000, 8, "tail_heavy")]
    L = Lists(specs, seed=1700)
    for cfg in (rz.TileConfig(512, 8, 16, 2, 1, 1), rz.TileConfig(256, 4, 16, 2, 2, 1),
                rz.TileConfig(512, 8, 16, 4, 2, 1), rz.TileConfig(256, 8, 1, 2, 1, 1)):
        b = ctx.batch(L.pairs(), config=cfg)
        for rep in range(2):
            announce("batch spread", cfg.as_tuple(), rep)
            L.restore()
            b.launch()
            L.check(oracle, f"spread {cfg.as_tuple()} launch {rep}")
        b.close()
    assert ctx.check_errors() == 0


def test_batch_bucketed_marks(ctx, oracle):
    """Batch variant 5: lists of 2^19 to 2^22 removals mark their tails in
    buckets (MARK items append entry runs, BUCKET items write bitmap words
    whole); smaller lists keep the atomic marks in the same batch.  First
    launches, a relaunch, a graph replay, RZ_PRESERVE_REMOVALS and re-planned
    one-shot calls, each list against the oracle; then a list that breaks the
    contract (a duplicate tail removal) is reported and the next batch on the
    context is exact (the bucket counts start from zero again)."""
    rng = np.random.default_rng(2525)
    specs = _random_specs(rng, 60, 60000)
    specs += [(1 << 22, 1 << 21, 4, "uniform"), ((1 << 21) + 7, 900_000, 16, "clustered"),
              (1 << 21, 600_000, 8, "tail_heavy"), (1 << 20, 1 << 19, 4, "uniform"), (1 << 20, 600_000, 16, "uniform"),
              ((1 << 22) + 3, (1 << 22) - 1, 4, "uniform"), (3_000_000, 140_000, 8, "tail_heavy")]
    L = Lists(specs, seed=2600)
    for cfg in (rz.TileConfig(512, 8, 0, 5, 1, 1), rz.TileConfig(256, 8, 0, 5, 2, 1), rz.TileConfig(512, 4, 0, 5, 2, 1)):
        b = ctx.batch(L.pairs(), config=cfg)
        for rep in range(2):
            announce("batch bucketed", cfg.as_tuple(), rep)
            L.restore()
            b.launch()
            L.check(oracle, f"bucketed {cfg.as_tuple()} launch {rep}")
        b.close()
    cfg = rz.TileConfig(512, 8, 0, 5, 1, 1)
    b = ctx.batch(L.pairs(), config=cfg, flags=rz.RZ_PRESERVE_REMOVALS)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        L.restore()
        torch.cuda.synchronize()
        b.capture(stream=stream)
        for rep in range(3):
            announce("batch bucketed replay", rep)
            L.restore()
            torch.cuda.synchronize()
            b.launch(stream=stream)
            stream.synchronize()
            L.check(oracle, f"bucketed replay {rep}")
            for d, h in zip(L.dev_R, L.host_R):
                assert torch.equal(d.cpu(), h)
    b.close()
    old = ctx.batch_config()
    ctx.set_batch_config(cfg)
    try:
        for seed in (2700, 2800):
            rng = np.random.default_rng(seed)
            specs2 = _random_specs(rng, 20, 40000) + [(1 << 21, int(rng.integers(1 << 19, 1 << 20)), 4, "uniform")]
            L2 = Lists(specs2, seed=seed)
            announce("batch bucketed one shot", seed)
            ctx.remove_batched(L2.pairs())
            L2.check(oracle, f"bucketed one-shot {seed}")
        assert ctx.check_errors() == 0
        # a duplicate tail removal in a bucketed list
        n, k = 1 << 21, 600_000
        A = W.unique_u32_payload(n, seed=2900).cuda()
        head = W.uniform_removals(n - k, k - 2, seed=2901)
        R = torch.cat([torch.tensor([n - 1, n - 1], dtype=torch.int32), head]).cuda()
        announce("batch bucketed mismatch")
        ctx.remove_batched([(A, R)])
        with pytest.raises(rz.RedzoneError) as e:
            ctx.check_errors()
        assert e.value.error_flags & rz.RZ_ERR_ORPHAN_MISMATCH
        L.restore()
        ctx.remove_batched(L.pairs())
        L.check(oracle, "bucketed after mismatch")
    finally:
        ctx.set_batch_config(old)
    assert ctx.check_errors() == 0
