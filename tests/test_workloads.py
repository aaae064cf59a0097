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


@pytest.mark.parametrize("n,frac", [(1000, 0.02), (50000, 0.1), (50001, 0.5), (4096, 0.25)])
@pytest.mark.parametrize("r", [100, -1, 0, 25, 50, 75])
def test_redzone_pct_lists(n, frac, r):
    """DataStruct.hxx:97-122: floor(k r / 100) removals in the red zone
    [n-k, n), listed after the green ones, all distinct and in range."""
    k = int(n * frac)
    R = W.redzone_pct_removals(n, k, r, seed=5).numpy()
    assert R.shape == (k,) and len(np.unique(R)) == k
    assert R.min() >= 0 and R.max() < n
    if r >= 0:
        red = (k * r) // 100
        tail = R >= n - k
        assert int(tail.sum()) == red
        assert not tail[: k - red].any() and tail[k - red:].all()
    assert np.array_equal(R, W.redzone_pct_removals(n, k, r, seed=5).numpy())


def test_redzone_pct_rejects_small_green_zone():
    with pytest.raises(ValueError):
        W.redzone_pct_removals(1000, 900, 0, seed=1)


@pytest.mark.parametrize("r", [0, 25, 50, 75, 100])
def test_redzone_pct_orphans_by_construction(oracle, r):
    """The orphan structure the red-zone share fixes: k_z = red_k; r = 0 has no
    orphans (every hole takes its own tail item), r = 100 has no holes below
    the red zone (the survivors stay where they are); in between the dest and
    source orphan counts are equal (paper Lemma 4.2)."""
    n, k = 1 << 16, 1 << 12
    A = np.arange(n, dtype=np.int32)
    R = W.redzone_pct_removals(n, k, r, seed=11).numpy()
    sz = n - k
    dead = np.zeros(k, bool)
    dead[R[R >= sz] - sz] = True
    dst = int(((R < sz) & dead).sum())
    src = int(((R >= sz) & ~dead).sum())
    assert dst == src
    want = A.copy()
    assert oracle.canonical(want, R.copy()) == 0
    if r == 0:
        assert dst == 0
        exp = A.copy()
        exp[R] = A[sz:]
        assert np.array_equal(want[:sz], exp[:sz])
    if r == 100:
        assert np.array_equal(want[:sz], A[:sz])


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_edge_payload_has_every_edge_word(esz):
    n = 20000
    P = W.edge_payload(esz, n, seed=3).numpy().reshape(n, -1).view(np.uint32)
    for word in W.EDGE_WORDS:
        assert (P[: n // 2] == word).any() and (P[n // 2:] == word).any()
    assert np.array_equal(P, W.edge_payload(esz, n, seed=3).numpy().reshape(n, -1).view(np.uint32))
