"""Seeded workload generators: distinct, in range, exact counts, deterministic,
and each pattern has its defining statistic."""
import numpy as np
import pytest
import torch

from redzone_amd import workloads as W


@pytest.mark.parametrize("n", [1, 2, 3, 17, 1000, 65536, 100003])
def test_permute_range_is_bijection(n):
    y = W.permute_range(torch.arange(n, dtype=torch.int64), n, seed=7).numpy()
    assert np.array_equal(np.sort(y), np.arange(n))


@pytest.mark.parametrize("pattern", ["uniform", "clustered", "tail_heavy"])
@pytest.mark.parametrize("n,frac", [(1000, 0.0), (1000, 0.01), (50000, 0.1), (50000, 0.5), (4096, 1.0), (20000, 0.9)])
def test_removals_distinct_in_range(pattern, n, frac):
    k = int(n * frac)
    R = W.removals(pattern, n, k, seed=3).numpy()
    assert R.dtype == np.int32 and R.shape == (k,)
    assert len(np.unique(R)) == k
    if k:
        assert R.min() >= 0 and R.max() < n
    R2 = W.removals(pattern, n, k, seed=3).numpy()
    assert np.array_equal(R, R2)


def test_uniform_tail_fraction():
    n, k = 1 << 20, 1 << 14
    R = W.uniform_removals(n, k, seed=1).numpy()
    kz = int((R >= n - k).sum())
    assert abs(kz - k * k / n) < 5 * np.sqrt(k * k / n)  # E[k_z] = k^2/n = 256


def test_tail_heavy_fraction():
    n = 1 << 18
    k = n // 100
    R = W.tail_heavy_removals(n, k, seed=2).numpy()
    # fraction of removals in the top 1% of positions: ~ 1 - 0.99^5 = 4.9%
    frac = (R >= n - k).mean()
    assert 0.035 < frac < 0.065
    # Beta(5,1) mean is 5/6
    assert abs(R.mean() / n - 5 / 6) < 0.02


def test_clustered_runs():
    n, k = 1 << 20, 1 << 15
    R = W.clustered_removals(n, k, seed=4, mean=256).numpy()
    steps = np.diff(R.astype(np.int64))
    run_breaks = int((steps != 1).sum()) + 1
    mean_run = k / run_breaks
    assert 150 < mean_run < 400


def test_payloads_unique_and_sized():
    n = 50000
    a = W.unique_u32_payload(n, 5).numpy()
    assert len(np.unique(a)) == n and (a < 0).any()  # high bit occurs
    b = W.u64_payload(n, 5).numpy()
    assert b.shape == (n, 2) and len(np.unique(b[:, 0])) == n
    c = W.rec16_payload(n, 5).numpy()
    assert c.shape == (n, 4) and len(np.unique(c[:, 0])) == n
    vals = c[:, 3].view(np.float32)
    assert abs(vals.mean()) < 0.05 and abs(vals.std() - 1) < 0.05
