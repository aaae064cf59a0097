"""Mixed streams: the generated lists are valid, the timing collection and the
trained dispatcher run, and every list processed by the batched + learned
path equals the oracle (canonical Redzone result, or the exact stable order
when the dispatcher picked the O(n) partition)."""
import numpy as np
import pytest
import torch

from redzone_amd import dispatch as D
from redzone_amd import mixed as MX
from tests.parity import announce, expected_canonical, first_mismatch

pytestmark = pytest.mark.gpu


def test_specs_and_lists_valid(ctx):
    specs = MX.stream_specs(30, seed=9, log2n=(4, 16))
    assert specs == MX.stream_specs(30, seed=9, log2n=(4, 16))
    for i, spec in enumerate(specs):
        n, k, esz, pattern = spec
        announce("mixed list", i, spec)
        A, R = MX.build_list(ctx, spec, 100 + i)
        assert A.shape[0] == n and R.shape[0] == k
        r = R.cpu().numpy()
        assert len(np.unique(r)) == k and (k == 0 or (r.min() >= 0 and r.max() < n))


def test_batched_learned_path_matches_oracle(ctx, oracle):
    specs = MX.stream_specs(60, seed=11, log2n=(8, 18))
    announce("collect")
    feats, times = MX.collect(ctx, specs[:30], seed=500, reps=1)
    assert np.all(times > 0)
    announce("train")
    disp = MX.train_model(ctx, feats, times, epochs=30)
    lists = [MX.build_list(ctx, s, 900 + 7 * i) for i, s in enumerate(specs)]
    host = [(A.cpu().numpy(), R.cpu().numpy()) for A, R in lists]
    announce("run batched learned")
    # a threshold inside the stream's k range, so both the batch and the per-list launches run
    ks = sorted(s[1] for s in specs)
    bmk = ks[len(ks) // 2]
    _, choice = MX.run_batched_learned(ctx, disp, lists, specs, batch_max_k=bmk)
    torch.cuda.synchronize()
    for i, ((A0, R0), (A, _)) in enumerate(zip(host, lists)):
        n, k = specs[i][0], specs[i][1]
        got = A.cpu().numpy()
        # lists below the batch threshold run in the (Redzone) batch whatever the choice
        if choice[i] == 3 and k >= bmk:
            bad, fb = oracle.check(A0, got, R0, mode=1)
            assert bad == 0, (i, specs[i], bad, fb)
        else:
            want = expected_canonical(oracle, A0, R0)
            m = first_mismatch(got[: n - k], want[: n - k])
            assert m < 0, (i, specs[i], m)


def test_plain_rule_path_matches_oracle(ctx, oracle):
    specs = MX.stream_specs(40, seed=12, log2n=(6, 16))
    lists = [MX.build_list(ctx, s, 300 + 7 * i) for i, s in enumerate(specs)]
    host = [(A.cpu().numpy(), R.cpu().numpy()) for A, R in lists]
    announce("run plain rule")
    MX.run_plain(ctx, lists, specs)
    for i, ((A0, R0), (A, _)) in enumerate(zip(host, lists)):
        n, k = specs[i][0], specs[i][1]
        got = A.cpu().numpy()
        if D.rule_choice(n, k) == 3:
            assert oracle.check(A0, got, R0, mode=1)[0] == 0
        else:
            want = expected_canonical(oracle, A0, R0)
            assert first_mismatch(got[: n - k], want[: n - k]) < 0


def test_remove_auto_matches_oracle_and_dispatcher(ctx, oracle):
    """rz_remove_auto: the method chosen in the library equals the Python
    dispatcher's choice for the same list, and the result equals the oracle
    (canonical Redzone, or the stable order for the partition).  Without
    parameters it follows the k < n/2 rule."""
    disp = D.Dispatcher.load(ctx)
    ctx.set_dispatch_params(disp.params.cpu().numpy())
    specs = MX.stream_specs(40, seed=13, log2n=(8, 18))
    specs += [(5000, 4000, 4, "uniform"), (5000, 10, 8, "clustered")]
    for i, spec in enumerate(specs):
        n, k = spec[0], spec[1]
        announce("auto", i, spec)
        A, R = MX.build_list(ctx, spec, 1300 + 7 * i)
        A0, R0 = A.cpu().numpy(), R.cpu().numpy()
        want_m = int(disp.choose(disp.features([(A, R)]))[0]) if 0 < k < n else 0
        new_len, m = ctx.remove_auto(A, R)
        torch.cuda.synchronize()
        assert new_len == n - k and m == want_m, (spec, m, want_m)
        got = A.cpu().numpy()
        if m == 3:
            assert oracle.check(A0, got, R0, mode=1)[0] == 0
        else:
            want = expected_canonical(oracle, A0, R0)
            assert first_mismatch(got[: n - k], want[: n - k]) < 0, (spec, m)
    ctx.set_dispatch_params(None)
    A, R = MX.build_list(ctx, (5000, 4000, 4, "uniform"), 77)
    _, m = ctx.remove_auto(A, R)
    assert m == 3
    A, R = MX.build_list(ctx, (5000, 100, 4, "uniform"), 78)
    _, m = ctx.remove_auto(A, R)
    assert m == 0
    ctx.set_dispatch_params(disp.params.cpu().numpy())
