"""Full-size parity of the tuned path: n = 2^28 unique uint32 payloads at every
point of the density / pattern sweep (k/n in {1e-4, 1e-3, 1e-2, 0.1, 0.5}
x {uniform, clustered runs of mean 256, tail-heavy Beta(5,1)}), through
rz_remove with no explicit launch shape, so each point runs the shape the
tuned table holds for its bucket (the one-launch fused form, the fused form and
the dense form with bucketed marks), and the dense form with atomic marks at
k/n = 0.1, named explicitly (no bucket holds it).  Also the reference's
red-zone-share lists (DataStruct.hxx:97-122) at r = 25 / 50 / 75 % on the
headline shape, the regime with the most orphans.

Bar: the survivors are array-identical to the canonical oracle and the device
error word is clear.
"""
import numpy as np
import pytest
import torch

from redzone_amd import workloads as W
from tests.parity import announce, expected_canonical, first_mismatch

pytestmark = pytest.mark.gpu

N = 1 << 28
FRACS = [1e-4, 1e-3, 1e-2, 0.1, 0.5]


def _run_full(ctx, oracle, A, R, what, config=None):
    n, k = A.shape[0], R.shape[0]
    shape = (config or ctx.table(4, k.bit_length() - 1)).as_tuple()
    announce(what, "k", k, "shape", shape)
    A_np = A.cpu().numpy()
    R_np = R.cpu().numpy()
    assert ctx.remove(A, R, config=config) == n - k
    torch.cuda.synchronize()
    assert ctx.check_errors() == 0
    got = A[: n - k].cpu().numpy()
    del A
    want = expected_canonical(oracle, A_np, R_np)
    m = first_mismatch(got, want[: n - k])
    assert m < 0, f"{what}: first mismatch at {m} (shape {shape})"


@pytest.mark.parametrize("pattern", ["uniform", "clustered", "tail_heavy"])
@pytest.mark.parametrize("frac", FRACS)
def test_config2_point_tuned(ctx, oracle, pattern, frac):
    k = int(N * frac)
    A = W.unique_u32_payload(N, seed=2026, device="cuda")
    R = W.removals(pattern, N, k, seed=2027, device="cuda")
    _run_full(ctx, oracle, A, R, f"config2 {pattern} {frac:g}")
    torch.cuda.empty_cache()


@pytest.mark.parametrize("pattern", ["uniform", "tail_heavy"])
def test_dense_atomic_marks_full_size(ctx, oracle, pattern):
    import redzone_amd as rz

    k = int(N * 0.1)
    A = W.unique_u32_payload(N, seed=2026, device="cuda")
    R = W.removals(pattern, N, k, seed=2027, device="cuda")
    _run_full(ctx, oracle, A, R, f"dense atomic marks {pattern}", config=rz.TileConfig(1024, 8, 1024, 5, 11, 0))
    torch.cuda.empty_cache()


@pytest.mark.parametrize("red_pct", [25, 50, 75])
def test_redzone_share_full_size(ctx, oracle, red_pct):
    k = N // 100
    A = W.unique_u32_payload(N, seed=31, device="cuda")
    R = W.redzone_pct_removals(N, k, red_pct, seed=32, device="cuda")
    _run_full(ctx, oracle, A, R, f"red zone share {red_pct}%")
    torch.cuda.empty_cache()
