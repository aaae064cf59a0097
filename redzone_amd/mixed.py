#!/usr/bin/env python3
"""Mixed streams of lists: training data for the learned dispatcher and the
mixed-stream evaluation (batched + learned dispatch vs plain per-list
launches with the k < n/2 rule).

Stream shape (one spec per list): n log-uniform in [2^12, 2^26], k/n
log-uniform in [1e-5, 0.9] (k = floor(f n)), items of 4, 8 or 16 bytes, and
removals uniform / clustered (geometric runs, mean 64) / tail-heavy
(Beta(5,1)) in equal thirds.  Lists are built on the device (rz_gen_* for
payloads and uniform removals, workloads.py torch code for the other two
patterns) and processed in resident batches.

    python -m redzone_amd.mixed train [--lists 4000] [--seed 1]   # writes the model
    python -m redzone_amd.mixed eval  [--lists 10000] [--seed 2]  # prints one JSON line
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PATTERNS = ("uniform", "clustered", "tail_heavy")


def stream_specs(count, seed, log2n=(12, 26), frac=(1e-5, 0.9), sizes=(4, 8, 16)):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(2 ** rng.uniform(*log2n))
        f = 10 ** rng.uniform(math.log10(frac[0]), math.log10(frac[1]))
        k = int(f * n)
        esz = int(sizes[int(rng.integers(len(sizes)))])
        specs.append((n, k, esz, PATTERNS[i % 3]))
    return specs


def build_list(ctx, spec, seed):
    """Device tensors (data, removals) of one spec; payload and removals seeded by `seed`."""
    import torch

    from . import workloads as W

    n, k, esz, pattern = spec
    shape = (n,) if esz == 4 else (n, esz // 4)
    A = torch.empty(shape, dtype=torch.int32, device="cuda")
    ctx.gen_payload(A, seed)
    if pattern == "uniform":
        R = torch.empty(k, dtype=torch.int32, device="cuda")
        ctx.gen_uniform_removals(R, n, seed + 1)
    elif pattern == "clustered":
        R = W.clustered_removals(n, k, seed + 1, mean=64, device="cuda")
    else:
        R = W.tail_heavy_removals(n, k, seed + 1, device="cuda")
    return A, R


def time_methods(ctx, A0, R0, methods, reps=3):
    """Median microseconds of each method on copies of (A0, R0), one list per launch."""
    import torch

    from .dispatch import run_method

    A = torch.empty_like(A0)
    R = torch.empty_like(R0)
    out = []
    for m in methods:
        ts = []
        for r in range(reps + 1):
            A.copy_(A0)
            R.copy_(R0)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            run_method(ctx, m, A, R)
            e1.record()
            e1.synchronize()
            if r:
                ts.append(e0.elapsed_time(e1) * 1e3)
        out.append(float(np.median(ts)))
    return np.array(out)


def batches(specs, budget_bytes):
    """Consecutive groups of specs whose data fits the budget."""
    group, used = [], 0
    for i, s in enumerate(specs):
        b = s[0] * s[2] + s[1] * 4
        if group and used + b > budget_bytes:
            yield group
            group, used = [], 0
        group.append(i)
        used += b
    if group:
        yield group


def collect(ctx, specs, seed, reps=3, budget=8 << 30, log=None):
    """Features [N, 12] and per-method times [N, 4] (us) of every spec."""
    import torch

    from .dispatch import METHODS, Dispatcher

    disp = Dispatcher(ctx)
    feats = np.zeros((len(specs), 12), np.float32)
    times = np.zeros((len(specs), len(METHODS)), np.float64)
    t0 = time.time()
    for g in batches(specs, budget):
        lists = [build_list(ctx, specs[i], seed + 7 * i) for i in g]
        feats[g] = disp.features(lists).cpu().numpy()
        for j, i in enumerate(g):
            times[i] = time_methods(ctx, lists[j][0], lists[j][1], range(len(METHODS)), reps)
        del lists
        torch.cuda.empty_cache()
        if log:
            log(f"collected {g[-1] + 1}/{len(specs)} lists ({time.time() - t0:.0f} s)")
    return feats, times


def train_model(ctx, feats, times, epochs=600, lr=3e-3, seed=0):
    import torch

    from .dispatch import Dispatcher

    X = torch.from_numpy(feats).cuda()
    T = torch.from_numpy(np.log2(np.maximum(times, 0.5)).astype(np.float32)).cuda()
    M = torch.ones_like(T)
    disp = Dispatcher(ctx, seed=seed)
    disp.train(X, T, M, epochs=epochs, lr=lr, batch=256, log_every=100)
    return disp


def _launch_each(ctx, rows, methods, stream):
    """One C-ABI launch per list (rows: rz_list_desc tuples), as a native host
    would issue them: rz_remove / rz_remove_config / rz_partition_on."""
    from . import RedzoneError
    from .dispatch import method_config

    lib = ctx._lib
    h = ctx.handle
    remove, remove_cfg, part = lib.rz_remove, lib.rz_remove_config, lib.rz_partition_on
    cfgs = {}
    for (d, r, n, k, e, _), m in zip(rows, methods):
        if m == 3:
            st = part(h, d, n, e, r, k, 0, stream)
        elif m == 0:
            st = remove(h, d, n, e, r, k, 0, stream)
        else:
            key = (m, e)
            if key not in cfgs:
                cfgs[key] = ctypes.byref(method_config(m, e))
            st = remove_cfg(h, d, n, e, r, k, 0, cfgs[key], stream)
        if st:
            raise RedzoneError(st, "per-list launch")


def _stream_arg(ctx):
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream(ctx.device).cuda_stream)


def _as_descs(ctx, lists):
    from . import ListDescs

    return lists if isinstance(lists, ListDescs) else ctx.descs(lists)


def run_plain(ctx, lists, specs=None):
    """Per-list launches, method by the k < n/2 rule; wall time (s) from first
    launch to the end.  lists: [(data, removals), ...] or a ListDescs (built
    before the clock starts, as a native caller holds its descriptors)."""
    import torch

    from .dispatch import rule_choice

    descs = _as_descs(ctx, lists)
    rows = descs.rows.tolist()
    methods = [rule_choice(n, k) for (_, _, n, k, _, _) in rows]
    s = _stream_arg(ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _launch_each(ctx, rows, methods, s)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


# Lists with fewer removals than this go into one persistent queue launch,
# whatever method the network prefers for them on their own: a list's own
# launch costs about 10 us, a list inside a batch about half a microsecond, and
# the queue moves large lists at the same random-write rate as their own
# launch.  Larger lists are launched on their own with the chosen method
# (measured on the 10,000-list stream: 2^20 175 ms, 2^22 170 ms, 2^24 190 ms).
BATCH_MAX_K = 1 << 22


def run_batched_learned(ctx, disp, lists, specs=None, batch_max_k=BATCH_MAX_K, batch_partition=False):
    """Lists below batch_max_k removals in one persistent launch of Redzone
    items; the others launched on their own with the method the MLP picks.
    The network's choice for the large lists is issued first, in one library
    call that returns at once (rz_dispatch_choose: features, forward pass and
    the copy back of the choices, in stream order); the batch is then planned
    and launched, and while the host plans, the device computes the choice, so
    that collecting it waits for nothing and the large lists queue up behind
    the batch.  With a trained network the whole sequence is one library call
    (rz_remove_auto_batched), as a native caller would make it; an untrained
    dispatcher (the k < n/2 rule) and batch_partition take the same steps from
    Python.  (Measured on the 10,000-list stream: deciding after the
    batch's launch left the device idle between the batch and the large
    lists, because the persistent batch holds every CU until it ends, and
    deciding first through separate feature and forward calls from Python
    cost more host time than the batch's plan.)
    With batch_partition the network also decides for every batched list
    before the batch is planned, and a batched list it sends to the partition
    runs the batch's partition items (the stable O(n) compaction).  That is
    slower on the mixed stream (measured: the batch 40.6 ms with Redzone items
    only, 50.5 ms with the network's partition picks, and the decision for
    every list adds 26 ms before the batch), because a partition list marks
    all k removals and streams all n items where Redzone items mark the tail's
    removals and write the holes; hence the default.  (The large lists' own
    launches were also tried on a second stream, beside the batch, through a
    second context: no gain, the two share the memory system.)  Wall time (s)
    of everything including the decision, the descriptor upload and the batch
    plan; returns (wall, method that ran per list).  lists: [(data,
    removals), ...] or a ListDescs."""
    import torch

    descs = _as_descs(ctx, lists)
    n, k = descs.rows["n"], descs.rows["k"]
    small = np.nonzero(k < batch_max_k)[0]
    large = np.nonzero(k >= batch_max_k)[0]
    choice = np.zeros(len(descs), np.int32)  # Redzone unless the network picks otherwise
    s = _stream_arg(ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    decide_small = batch_partition and disp.trained and len(small) > 0
    if disp.trained and not decide_small:  # the whole sequence in one library call
        from . import RedzoneError
        from .dispatch import _ptr

        st = ctx._lib.rz_remove_auto_batched(ctx.handle, _ptr(disp.params), descs.as_ctypes(), len(descs),
                                             int(batch_max_k), 0,
                                             choice.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), s)
        if st:
            raise RedzoneError(st, "rz_remove_auto_batched")
        torch.cuda.synchronize()
        return time.perf_counter() - t0, choice
    C = None
    if decide_small:  # every list's method before the batch is planned
        disp.choose_async(descs)
        C = disp.collect()
    elif len(large):
        disp.choose_async(descs.subset(large))  # computed on the device while the batch is planned
    if len(small):
        sub = descs.subset(small)
        if decide_small:
            sub.rows["method"] = np.where(C[small] == 3, 3, 0)
            choice[small] = sub.rows["method"]
        ctx.remove_batched(sub)
    if len(large):
        choice[large] = C[large] if decide_small else disp.collect()
        _launch_each(ctx, descs.rows[large].tolist(), [int(choice[i]) for i in large], s)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, choice


def evaluate(ctx, disp, specs, seed, reps=3, budget=8 << 30, log=None, sweep=(1 << 20, 1 << 24), ctx_untuned=None):
    """Mixed-stream evaluation; returns a dict of totals and regrets.  `sweep`
    lists further batch thresholds whose batched wall time is also measured.
    With ctx_untuned (a Context without the tuned table: every list runs the
    fixed default shape, three kernels) the plain per-list launches are also
    timed on it."""
    import torch

    from .dispatch import METHODS, rule_choice

    tot = {"plain_rule_wall_s": 0.0, "batched_learned_wall_s": 0.0, "oracle_best_us": 0.0, "learned_us": 0.0,
           "rule_us": 0.0, "redzone_only_us": 0.0}
    if ctx_untuned is not None:
        tot["plain_untuned_wall_s"] = 0.0
    swept = {int(b): 0.0 for b in sweep}
    picks = np.zeros(4, np.int64)
    per_list = np.zeros((len(specs), 4), np.float64)
    chosen = np.zeros(len(specs), np.int32)
    t0 = time.time()
    for g in batches(specs, budget):
        sub = [specs[i] for i in g]
        pristine = [build_list(ctx, specs[i], seed + 7 * i) for i in g]
        X = disp.features(pristine)
        learned = disp.choose(X)
        for j, i in enumerate(g):
            t = time_methods(ctx, pristine[j][0], pristine[j][1], range(4), reps)
            per_list[i] = t
            chosen[i] = learned[j]
            n, k = specs[i][0], specs[i][1]
            tot["oracle_best_us"] += t.min()
            tot["learned_us"] += t[learned[j]]
            tot["rule_us"] += t[rule_choice(n, k)]
            tot["redzone_only_us"] += t[0]
            picks[learned[j]] += 1
        work = [(torch.empty_like(A), torch.empty_like(R)) for A, R in pristine]
        wdescs = ctx.descs(work)

        def restore():
            for (a, r), (a0, r0) in zip(work, pristine):
                a.copy_(a0)
                r.copy_(r0)
            torch.cuda.synchronize()

        # every path runs once untimed first, so that its scratch buffers have
        # grown to this group's sizes before the timed run
        def timed(fn):
            restore()
            fn()
            restore()
            return fn()

        tot["plain_rule_wall_s"] += timed(lambda: run_plain(ctx, wdescs))
        if ctx_untuned is not None:
            udescs = ctx_untuned.descs(work)
            tot["plain_untuned_wall_s"] += timed(lambda: run_plain(ctx_untuned, udescs))
        tot["batched_learned_wall_s"] += timed(lambda: run_batched_learned(ctx, disp, wdescs)[0])
        for b in swept:
            swept[b] += timed(lambda: run_batched_learned(ctx, disp, wdescs, batch_max_k=b)[0])
        del pristine, work
        torch.cuda.empty_cache()
        if log:
            log(f"evaluated {g[-1] + 1}/{len(specs)} lists ({time.time() - t0:.0f} s)")
    tot["learned_regret"] = tot["learned_us"] / tot["oracle_best_us"] - 1
    tot["rule_regret"] = tot["rule_us"] / tot["oracle_best_us"] - 1
    tot["batched_learned_vs_plain"] = tot["plain_rule_wall_s"] / tot["batched_learned_wall_s"]
    if ctx_untuned is not None:
        tot["batched_learned_vs_plain_untuned"] = tot["plain_untuned_wall_s"] / tot["batched_learned_wall_s"]
    tot["batch_max_k"] = BATCH_MAX_K
    tot["batched_wall_s_by_batch_max_k"] = {str(b): w for b, w in swept.items()}
    tot["picks"] = {m: int(c) for m, c in zip(METHODS, picks)}
    tot["_per_list_us"] = per_list
    tot["_chosen"] = chosen
    return tot


def main():
    import redzone_amd as rz

    from .dispatch import MODEL_PATH, Dispatcher

    p = argparse.ArgumentParser()
    p.add_argument("mode", choices=["train", "eval"])
    p.add_argument("--lists", type=int, default=None)
    p.add_argument("--seed", type=int, default=None)
    p.add_argument("--reps", type=int, default=3)
    p.add_argument("--model", default=MODEL_PATH)
    p.add_argument("--epochs", type=int, default=600)
    p.add_argument("--details", default=None, help="eval: write per-list times and choices to this .npz")
    args = p.parse_args()

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    ctx = rz.Context(0)
    if args.mode == "train":
        count, seed = args.lists or 4000, 1 if args.seed is None else args.seed
        specs = stream_specs(count, seed)
        feats, times = collect(ctx, specs, seed * 1000003, args.reps, log=log)
        disp = train_model(ctx, feats, times, epochs=args.epochs)
        disp.save(args.model, meta={"lists": count, "seed": seed, "train_log": disp.train_log})
        X = feats
        pred, _ = disp.predict(__import__("torch").from_numpy(X).cuda())
        choice = np.argmin(pred.cpu().numpy(), axis=1)
        best = times.min(1).sum()
        print(json.dumps({"mode": "train", "lists": count, "seed": seed, "train_log": disp.train_log,
                          "train_regret": float(times[np.arange(len(specs)), choice].sum() / best - 1),
                          "model": args.model}), flush=True)
    else:
        count, seed = args.lists or 10000, 2 if args.seed is None else args.seed
        disp = Dispatcher.load(ctx, args.model)
        specs = stream_specs(count, seed)
        res = evaluate(ctx, disp, specs, seed * 1000003, args.reps, log=log,
                       ctx_untuned=rz.Context(ctx.device, load_tuned=False, load_dispatcher=False))
        per_list, chosen = res.pop("_per_list_us"), res.pop("_chosen")
        if args.details:
            np.savez(args.details, specs=np.array(specs, dtype=object).astype(str), times_us=per_list, chosen=chosen)
        res.update({"mode": "eval", "lists": count, "seed": seed})
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
