"""Device-detected failures reach the caller.

The kernels check two things the host cannot see: that the dest and source
orphan totals agree (the reference asserts the same equality in debug builds,
redzone.hxx:905) and that every bounded wait on other workgroups ended in
time.  A failed check sets a bit of the context's error word, which
rz_check_errors (and every call with RZ_STATS or RZ_VALIDATE) reads, clears and
turns into RZ_EDEVICE, after resetting the scratch the kernels keep zero.

The wait give-ups are forced from the host through the test hook
rz_debug_set_waits (a short poll cap, and a stall flag that makes the waits of
the fused forms see their condition as unmet); every kernel still runs to its
end inside its buffers.
"""
import pytest
import torch

import redzone_amd as rz
from redzone_amd import workloads as W
from tests.parity import announce, gpu_remove_parity

pytestmark = pytest.mark.gpu


def _stalled_call(ctx, dA, dR, cfg, flags=0):
    ctx.debug_set_waits(spin_cap=16, stall=True)
    try:
        return ctx.remove(dA, dR, config=cfg, flags=flags)
    finally:
        torch.cuda.synchronize()
        ctx.debug_set_waits(0, False)


@pytest.mark.parametrize("variant", [4, 6])
def test_forced_wait_giveup_reaches_python(ctx, oracle, variant):
    """Fused forms with every look-back (and, in the one-launch form, the wait
    for the marks) forced to give up: rz_check_errors reports RZ_EDEVICE with
    the matching bits, and the next call on the same context is exact again."""
    n, k = 1 << 20, (1 << 20) // 50
    A = W.unique_u32_payload(n, seed=5)
    R = W.uniform_removals(n, k, seed=6)
    cfg = rz.TileConfig(256, 4, 64, variant, 0, 0)  # 64 slices: 63 of them look back
    announce("forced give-up", variant)
    assert ctx.check_errors() == 0
    dA, dR = A.cuda(), R.cuda()
    _stalled_call(ctx, dA, dR, cfg)
    with pytest.raises(rz.RedzoneError) as e:
        ctx.check_errors()
    assert e.value.status == rz.RZ_EDEVICE
    assert e.value.error_flags & rz.RZ_ERR_LOOKBACK_TIMEOUT
    if variant == 6:
        assert e.value.error_flags & rz.RZ_ERR_WAIT_TIMEOUT
    assert ctx.check_errors() == 0, "the check clears the word"
    announce("after give-up", variant)
    gpu_remove_parity(ctx, oracle, A, R, config=cfg)


def test_stats_call_returns_device_error(ctx, oracle):
    """A call with RZ_STATS returns the failed check itself (no separate check)."""
    n, k = 1 << 20, (1 << 20) // 50
    A = W.unique_u32_payload(n, seed=7)
    R = W.uniform_removals(n, k, seed=8)
    cfg = rz.TileConfig(256, 4, 64, 4, 0, 0)
    announce("stats give-up")
    dA, dR = A.cuda(), R.cuda()
    with pytest.raises(rz.RedzoneError) as e:
        _stalled_call(ctx, dA, dR, cfg, flags=rz.RZ_STATS)
    assert e.value.status == rz.RZ_EDEVICE
    assert ctx.last_stats().error_flags & rz.RZ_ERR_LOOKBACK_TIMEOUT
    gpu_remove_parity(ctx, oracle, A, R, config=cfg, flags=rz.RZ_STATS)


@pytest.mark.parametrize("variant", [0, 3, 4, 5, 6])
def test_orphan_mismatch_is_reported(ctx, variant):
    """A removal list that breaks the contract (the last index listed twice, not
    validated): the orphan totals differ, which every form detects; the call
    with RZ_STATS returns RZ_EDEVICE with RZ_ERR_ORPHAN_MISMATCH."""
    n, k = 4096, 64
    A = W.unique_u32_payload(n, seed=9).cuda()
    head = W.uniform_removals(n - k, k - 2, seed=10)
    R = torch.cat([torch.tensor([n - 1, n - 1], dtype=torch.int32), head]).cuda()
    cfg = rz.TileConfig(256, 4, 4, variant, 0, 0)
    announce("mismatch", variant)
    with pytest.raises(rz.RedzoneError) as e:
        ctx.remove(A, R, config=cfg, flags=rz.RZ_STATS)
    assert e.value.status == rz.RZ_EDEVICE
    assert ctx.last_stats().error_flags & rz.RZ_ERR_ORPHAN_MISMATCH
    assert ctx.check_errors() == 0
    # RZ_VALIDATE rejects the same list before any kernel runs
    with pytest.raises(rz.RedzoneError) as e:
        ctx.remove(A, R, config=cfg, flags=rz.RZ_VALIDATE)
    assert e.value.status == 3


def test_refused_shape_runs_nothing(ctx, oracle):
    """A launch shape the library cannot run is refused before any kernel: no
    data touched, and no tail mark left behind for the next call (the dense
    form's 1024 x 8 staging does not fit LDS for 16-byte items; a workgroup
    size or item count with no kernel; an unknown variant)."""
    n, k = (1 << 20) + 3, 40_000
    A = W.payload(16, n, seed=12)
    R = W.uniform_removals(n, k, seed=13)
    dA, dR = A.cuda(), R.cuda()
    for bad in (rz.TileConfig(1024, 8, 1024, 5, 11, 0), rz.TileConfig(128, 4, 512, 4, 0, 0),
                rz.TileConfig(256, 3, 512, 0, 0, 0), rz.TileConfig(256, 4, 512, 9, 0, 0)):
        announce("refused shape", bad)
        with pytest.raises(rz.RedzoneError) as e:
            ctx.remove(dA, dR, config=bad)
        assert e.value.status == 1
    torch.cuda.synchronize()
    assert torch.equal(dA.cpu(), A) and torch.equal(dR.cpu(), R)
    assert ctx.check_errors() == 0
    # the three-kernel form reads the tail bitmap: a mark left behind would show here
    gpu_remove_parity(ctx, oracle, A, R, config=rz.TileConfig(256, 4, 64, 0, 0, 0))
