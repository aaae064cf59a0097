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
        # the choice is the method that ran: the partition (in the batch or on its own) or Redzone
        if choice[i] == 3:
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


def _check_list(oracle, spec, A0, R0, got, partition):
    n, k = spec[0], spec[1]
    if partition:
        bad, fb = oracle.check(A0, got, R0, mode=1)
        assert bad == 0, (spec, bad, fb)
    else:
        want = expected_canonical(oracle, A0, R0)
        m = first_mismatch(got[: n - k], want[: n - k])
        assert m < 0, (spec, m)


# A seeded sample of the mixed stream at its stated shape (n log-uniform in
# [2^12, 2^26], k/n log-uniform in [1e-5, 0.9], 4/8/16-byte items, three
# removal patterns), plus its largest corner: lists of about 2^26 items of
# every size, the largest 16-byte ones 1 GiB.
def _config5_sample():
    specs = MX.stream_specs(240, seed=17)
    specs += [((1 << 26) - 1, int(((1 << 26) - 1) * f), esz, pat)
              for esz, f, pat in [(16, 0.3, "uniform"), (16, 0.02, "tail_heavy"), (8, 0.9, "clustered"),
                                  (4, 0.07, "uniform"), (8, 1e-5, "uniform")]]
    return specs


def test_config5_sample_batched_learned(ctx, oracle):
    """Every list of the sample through run_batched_learned (the batch below
    2^22 removals, the network's method above), in resident groups, each list
    checked against the oracle; then the same with the network's partition
    picks run as the batch's partition items."""
    disp = D.Dispatcher.load(ctx)
    specs = _config5_sample()
    assert {s[2] for s in specs} == {4, 8, 16} and max(s[0] for s in specs) >= (1 << 26) - 1
    ran = {0: 0, 1: 0, 2: 0, 3: 0}
    for g in MX.batches(specs, 6 << 30):
        announce("config5 group", g[0], g[-1])
        for batch_partition in (False, True):
            lists = [MX.build_list(ctx, specs[i], 5000 + 7 * i) for i in g]
            host = [(A.cpu().numpy(), R.cpu().numpy()) for A, R in lists]
            _, choice = MX.run_batched_learned(ctx, disp, lists, batch_partition=batch_partition)
            torch.cuda.synchronize()
            assert ctx.check_errors() == 0
            for j, i in enumerate(g):
                m = int(choice[j])  # the method that ran (batched lists: Redzone, or partition items)
                if not batch_partition:
                    ran[m] += 1
                _check_list(oracle, specs[i], host[j][0], host[j][1], lists[j][0].cpu().numpy(), m == 3)
            del lists, host
            torch.cuda.empty_cache()
    assert ran[0] > 100 and sum(ran.values()) == len(specs)


def test_config5_sample_remove_auto(ctx, oracle):
    """The same sample list by list through rz_remove_auto (the network inside
    the library), each checked against the oracle; the network sends some
    lists to Redzone and some to the partition."""
    disp = D.Dispatcher.load(ctx)
    ctx.set_dispatch_params(disp.params.cpu().numpy())
    specs = _config5_sample()
    used = set()
    for i, spec in enumerate(specs):
        announce("config5 auto", i, spec)
        A, R = MX.build_list(ctx, spec, 5000 + 7 * i)
        A0, R0 = A.cpu().numpy(), R.cpu().numpy()
        new_len, m = ctx.remove_auto(A, R)
        torch.cuda.synchronize()
        assert new_len == spec[0] - spec[1]
        assert ctx.check_errors() == 0
        used.add(m)
        _check_list(oracle, spec, A0, R0, A.cpu().numpy(), m == 3)
        del A, R
    assert {0, 3} <= used, used


def test_learned_dispatch_gain(ctx):
    """What the trained network gains over the method without it: on a
    held-out sample of the mixed stream (300 lists of the stated shape, seed 3;
    the committed model was trained on seed 1), every method is timed on every
    list, and the total time of the network's picks is compared with the
    measured per-list best and with the k < n/2 rule (README.md:6)."""
    disp = D.Dispatcher.load(ctx)
    specs = MX.stream_specs(300, seed=3)
    best = learned = rule = 0.0
    for g in MX.batches(specs, 6 << 30):
        announce("gain group", g[0], g[-1])
        lists = [MX.build_list(ctx, specs[i], 9000 + 7 * i) for i in g]
        pick = disp.choose(disp.features(lists))
        for j, i in enumerate(g):
            n, k = specs[i][0], specs[i][1]
            t = MX.time_methods(ctx, lists[j][0], lists[j][1], range(4), reps=5)
            best += t.min()
            learned += t[int(pick[j])]
            rule += t[D.rule_choice(n, k)]
        del lists
        torch.cuda.empty_cache()
    learned_regret, rule_regret = learned / best - 1, rule / best - 1
    print(f"held-out sample: learned regret {learned_regret:.4f}, rule regret {rule_regret:.4f}, "
          f"learned saves {1 - learned / rule:.4f} of the rule's time", flush=True)
    assert learned_regret <= 0.05
    assert learned_regret < rule_regret
