"""Batches whose descriptors send some lists to the partition (method
RZ_METHOD_PARTITION): those lists run PMARK and PTILE items, the stable O(n)
compaction inside the persistent launch, beside Redzone lists.  A partition
list must equal the oracle's stable removal (numpy.delete order) exactly; a
Redzone list the canonical oracle, on first launches, relaunches, graph
replays and re-planned one-shot calls."""
import numpy as np
import pytest
import torch

import redzone_amd as rz
from tests.parity import announce
from tests.test_gpu_batched import Lists, _random_specs

pytestmark = pytest.mark.gpu


def _stable(oracle, A, R):
    want = A.copy()
    st, _ = oracle.remove_stable(want, R)
    assert st == oracle.OK
    return want


class Mixed(Lists):
    """Lists with a method per list (3 = partition, 0 = Redzone)."""

    def __init__(self, specs, seed, methods, oracle):
        super().__init__(specs, seed)
        self.methods = np.asarray(methods, np.uint32)
        for i, m in enumerate(self.methods):
            if m == rz.RZ_METHOD_PARTITION:
                self.want[i] = _stable(oracle, self.host_A[i].numpy(), self.host_R[i].numpy())

    def descs(self, ctx):
        d = ctx.descs(self.pairs())
        d.rows["method"] = self.methods
        return d


def _methods(specs, rng):
    # the dense lists and a random share of the others to the partition
    return [3 if (k > 0.3 * n or rng.random() < 0.3) else 0 for n, k, _, _ in specs]


def test_partition_items_small_lists(ctx, oracle):
    rng = np.random.default_rng(707)
    specs = _random_specs(rng, 300, 60000)
    L = Mixed(specs, 9000, _methods(specs, rng), oracle)
    for cfg in (None, rz.TileConfig(256, 4, 0, 2, 1, 1), rz.TileConfig(512, 8, 0, 4, 1, 1),
                rz.TileConfig(512, 8, 1, 2, 1, 1)):
        announce("partition items small", cfg.as_tuple() if cfg else None)
        L.restore()
        b = ctx.batch(L.descs(ctx), config=cfg)
        b.launch()
        L.check(oracle, f"partition items {cfg}")
        b.close()
    assert ctx.check_errors() == 0


def test_partition_items_large_lists(ctx, oracle):
    """Lists of up to 2^23 items in partition tiles (hundreds of tiles, so the
    look-back spans several windows) beside large Redzone lists."""
    specs = [(1 << 23, 1 << 22, 4, "uniform"), ((1 << 22) + 5, 3_000_000, 8, "clustered"),
             (3_000_001, 100, 16, "tail_heavy"), (1 << 22, 1 << 20, 4, "uniform"), (300_000, 250_000, 16, "uniform"),
             (70_000, 69_999, 8, "uniform"), (5000, 0, 4, "uniform"), (4096, 4096, 4, "uniform")]
    methods = [3, 3, 3, 0, 3, 3, 3, 3]
    L = Mixed(specs, 9100, methods, oracle)
    for cfg in (None, rz.TileConfig(256, 8, 0, 2, 2, 1)):
        announce("partition items large", cfg.as_tuple() if cfg else None)
        L.restore()
        b = ctx.batch(L.descs(ctx), config=cfg)
        b.launch()
        L.check(oracle, "partition large")
        b.close()
    assert ctx.check_errors() == 0


def test_partition_items_relaunch_graph_and_preserve(ctx, oracle):
    rng = np.random.default_rng(808)
    specs = _random_specs(rng, 120, 200000)
    L = Mixed(specs, 9200, _methods(specs, rng), oracle)
    b = ctx.batch(L.descs(ctx), flags=rz.RZ_PRESERVE_REMOVALS)
    for rep in range(2):
        announce("partition relaunch", rep)
        L.restore()
        b.launch()
        L.check(oracle, f"partition relaunch {rep}")
    for d, h in zip(L.dev_R, L.host_R):
        assert torch.equal(d.cpu(), h)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        L.restore()
        torch.cuda.synchronize()
        b.capture(stream=stream)
        for rep in range(3):
            announce("partition graph replay", rep)
            L.restore()
            torch.cuda.synchronize()
            b.launch(stream=stream)
            stream.synchronize()
            L.check(oracle, f"partition replay {rep}")
    b.close()
    assert ctx.check_errors() == 0


def test_partition_items_replanned(ctx, oracle):
    """rz_remove_batched re-plans every call: two list sets in a row, the
    second with other lists where the first one's tile status words were."""
    for seed in (9300, 9400):
        rng = np.random.default_rng(seed)
        specs = _random_specs(rng, 64, 150000)
        L = Mixed(specs, seed, _methods(specs, rng), oracle)
        announce("partition one shot", seed)
        ctx.remove_batched(L.descs(ctx))
        L.check(oracle, f"partition one-shot {seed}")
    assert ctx.check_errors() == 0
