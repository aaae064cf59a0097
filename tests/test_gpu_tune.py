"""Device generators (bit-exact against redzone_amd/workloads.py) and the
autotuner (installs the fastest shape; the installed shape gives the oracle's
result; the table survives a save/load round trip)."""
import os
import tempfile

import numpy as np
import pytest
import torch

import redzone_amd as rz
from redzone_amd import workloads as W
from tests.parity import announce, gpu_remove_parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,k,seed", [(1, 1, 3), (2, 1, 4), (1000, 300, 5), (1 << 20, 1 << 14, 6),
                                      ((1 << 20) + 17, 500000, 7), (1 << 24, 1 << 22, 8)])
def test_gen_uniform_removals_matches_python(ctx, n, k, seed):
    announce("gen removals", n, k, seed)
    R = torch.empty(k, dtype=torch.int32, device="cuda")
    ctx.gen_uniform_removals(R, n, seed)
    torch.cuda.synchronize()
    want = W.uniform_removals(n, k, seed)
    assert torch.equal(R.cpu(), want)


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_gen_payload_matches_python(ctx, esz):
    n = (1 << 18) + 3
    announce("gen payload", esz)
    shape = (n,) if esz == 4 else (n, esz // 4)
    A = torch.empty(shape, dtype=torch.int32, device="cuda")
    ctx.gen_payload(A, 42)
    torch.cuda.synchronize()
    want = W.payload(esz, n, 42)
    got = A.cpu()
    if esz == 16:
        assert torch.equal(got[:, :3], want[:, :3])  # the fourth word is a hash on the device, a float on the host
    else:
        assert torch.equal(got, want)


def test_tune_installs_best_and_stays_exact(oracle):
    ctx = rz.Context(0, load_tuned=False)
    n, k = 1 << 20, 1 << 14
    announce("tune", n, k)
    best, best_ms, results = ctx.tune(4, n, k, reps=2, cold=False)
    assert best_ms > 0
    timed = [ms for _, ms in results if ms > 0]
    assert min(timed) == pytest.approx(best_ms)
    assert ctx.table(4, 14).as_tuple() == best.as_tuple()
    A = W.unique_u32_payload(n, seed=3)
    R = W.uniform_removals(n, k, seed=4)
    gpu_remove_parity(ctx, oracle, A, R)  # default shape = the installed one
    ctx.close()


def test_table_round_trip():
    ctx = rz.Context(0, load_tuned=False)
    cfg = rz.TileConfig(512, 16, 333, 3, 0, 1)
    ctx.set_table(8, 21, cfg)
    bcfg = rz.TileConfig(256, 8, 0, 2, 4, 2)
    ctx.set_batch_config(bcfg)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "table.txt")
        ctx.save_table(path)
        ctx2 = rz.Context(0, load_tuned=False)
        assert ctx2.table(8, 21).as_tuple() != cfg.as_tuple()
        ctx2.load_table(path)
        assert ctx2.table(8, 21).as_tuple() == cfg.as_tuple()
        assert ctx2.batch_config().as_tuple() == bcfg.as_tuple()
        ctx2.close()
    ctx.close()


def test_tune_batch():
    ctx = rz.Context(0, load_tuned=False)
    announce("tune batch")
    best, best_ms, results = ctx.tune_batch(4, 64, 1 << 8, 1 << 14, 0.3, reps=2)
    assert best_ms > 0 and best.variant in (2, 4, 5)  # MOVE/PAIR items, fused items, bucketed marks
    assert {c.variant for c, _ in results} == {2, 4, 5}
    assert all(t > 0 for _, t in results), "a candidate was rejected (its output differed from the reference)"
    assert ctx.batch_config().as_tuple() == best.as_tuple()
    ctx.close()


def test_committed_table_loads():
    from redzone_amd import _native

    if not os.path.exists(_native.TUNED_TABLE):
        pytest.skip("no committed table")
    ctx = rz.Context(0)
    assert ctx.tuned_table == _native.TUNED_TABLE
    ctx.close()
