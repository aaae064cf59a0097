#!/usr/bin/env python3
"""Timing sweeps of the Redzone path on one device, one JSON line per point.

    python -m redzone_amd.sweep [--config 2|3|gather|partition|all] [--reps 5]

Config 2: n = 2^28 unique uint32 payloads, k/n in {1e-4, 1e-3, 1e-2, 0.1, 0.5},
removals uniform / clustered (geometric runs, mean 256) / tail-heavy (Beta(5,1)
positions); each point times rz_remove with the tuned launch shape, the dense
form, and the O(n) stable partition on the same list.

Config 3: n = 2^27 16-byte records {uint64 key, uint32 tag, float32 value},
k = n/100 uniform removals, the removal list once randomly permuted and once
sorted ascending (the sorted list is also timed with RZ_SORTED).

Gather: n = 2^28 uint32, a sorted uniform preserve list of p = n * keep
positions (keep in {0.001, 0.01, 0.1, 0.5, 0.99}) timed with the in-place
gather, and its complement as a removal list timed with Redzone and the O(n)
partition.

Partition: the O(n) stable partition on n = 2^28 uint32 at k/n = 0.3% and
50% uniform (the two mark forms), beside Redzone on the same list.

Every call starts from pristine copies of the list and the removal list, with
a 1 GiB read in between so that the caches hold no useful lines (cold
timing); a point's time is the median over `reps` calls, from events around
the call on the launch stream.  Correctness is the parity tests' business;
this tool only times.
"""
import argparse
import json
import os
import statistics
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _timer(torch):
    scrub = torch.ones((1 << 30) // 8, dtype=torch.int64, device="cuda")

    def time_call(restore, call, reps):
        ts = []
        for r in range(reps + 1):
            restore()
            scrub.sum()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            call()
            e1.record()
            e1.synchronize()
            if r:
                ts.append(e0.elapsed_time(e1))
        return statistics.median(ts)

    return time_call


def dense_config(rz, elem_bytes):
    return rz.TileConfig(1024, 4 if elem_bytes == 16 else 8, 1024, 5, 11, 2)


def sweep_config2(ctx, rz, W, torch, reps, emit):
    n = 1 << 28
    time_call = _timer(torch)
    A0 = W.payload(4, n, seed=2026, device="cuda")
    A = torch.empty_like(A0)
    for pattern in ("uniform", "clustered", "tail_heavy"):
        for frac in (1e-4, 1e-3, 1e-2, 0.1, 0.5):
            k = int(n * frac)
            kw = {"mean": 256} if pattern == "clustered" else {}
            R0 = W.removals(pattern, n, k, seed=2027, device="cuda", **kw)
            R = torch.empty_like(R0)

            def restore():
                A.copy_(A0)
                R.copy_(R0)

            t_rz = time_call(restore, lambda: ctx.remove(A, R), reps)
            t_dense = time_call(restore, lambda: ctx.remove(A, R, config=dense_config(rz, 4)), reps)
            t_on = time_call(restore, lambda: ctx.partition(A, R), max(2, reps // 2))
            shape = ctx.table(4, k.bit_length() - 1).as_tuple()
            emit({"config": 2, "n": n, "elem_bytes": 4, "pattern": pattern, "frac": frac, "k": k,
                  "redzone_ms": round(t_rz, 4), "tuned_shape": shape, "dense_ms": round(t_dense, 4),
                  "partition_ms": round(t_on, 4), "redzone_Gitems_s": round(n / (t_rz * 1e-3) / 1e9, 1),
                  "partition_over_redzone": round(t_on / min(t_rz, t_dense), 2)})
            del R0, R
            torch.cuda.empty_cache()


def sweep_config3(ctx, rz, W, torch, reps, emit):
    n = 1 << 27
    k = n // 100
    time_call = _timer(torch)
    A0 = W.payload(16, n, seed=2026, device="cuda")
    A = torch.empty_like(A0)
    Rp = W.uniform_removals(n, k, seed=2029, device="cuda")
    Rs = torch.sort(Rp).values.contiguous()
    for order, R0 in (("permuted", Rp), ("sorted", Rs)):
        R = torch.empty_like(R0)

        def restore():
            A.copy_(A0)
            R.copy_(R0)

        res = {"config": 3, "n": n, "elem_bytes": 16, "k": k, "order": order,
               "redzone_ms": round(time_call(restore, lambda: ctx.remove(A, R), reps), 4)}
        if order == "sorted":
            res["redzone_sorted_flag_ms"] = round(
                time_call(restore, lambda: ctx.remove(A, R, flags=rz.RZ_SORTED), reps), 4)
        res["partition_ms"] = round(time_call(restore, lambda: ctx.partition(A, R), max(2, reps // 2)), 4)
        moved = 16 * (2 * k)  # tail reads and hole writes, the record bytes a removal moves at most
        res["moved_GBps"] = round(moved / (min(res["redzone_ms"], res.get("redzone_sorted_flag_ms", 1e9)) * 1e-3) / 1e9,
                                  1)
        emit(res)


def sweep_gather(ctx, rz, W, torch, reps, emit):
    """The preserve-list gather against Redzone and the O(n) partition on the
    same split: P = a sorted uniform sample of p positions, R = its complement."""
    n = 1 << 28
    time_call = _timer(torch)
    A0 = W.payload(4, n, seed=2026, device="cuda")
    A = torch.empty_like(A0)
    for keep in (0.001, 0.01, 0.1, 0.5, 0.99):
        p = int(n * keep)
        Rm = W.uniform_removals(n, n - p, seed=2030, device="cuda")  # the removals, unsorted
        mask = torch.ones(n, dtype=torch.bool, device="cuda")
        mask[Rm.long()] = False
        P = torch.nonzero(mask).flatten().to(torch.int32)
        del mask
        R = torch.empty_like(Rm)

        def restore():
            A.copy_(A0)
            R.copy_(Rm)

        res = {"config": "gather", "n": n, "p": p, "keep": keep,
               "gather_ms": round(time_call(restore, lambda: ctx.gather_preserve(A, P, n=n), reps), 4),
               "redzone_ms": round(time_call(restore, lambda: ctx.remove(A, R), reps), 4),
               "partition_ms": round(time_call(restore, lambda: ctx.partition(A, R), max(2, reps // 2)), 4)}
        res["gather_GBps"] = round((4 * p + 4 * p + 4 * p) / (res["gather_ms"] * 1e-3) / 1e9, 1)
        emit(res)
        del Rm, R, P
        torch.cuda.empty_cache()


def sweep_partition(ctx, rz, W, torch, reps, emit):
    """The O(n) stable partition alone on n = 2^28 uint32 with uniform removals
    at k/n = 0.3% and 50% (both with bucketed marks: 2^19 removals and more),
    beside Redzone."""
    n = 1 << 28
    time_call = _timer(torch)
    A0 = W.payload(4, n, seed=2026, device="cuda")
    A = torch.empty_like(A0)
    for frac in (0.003, 0.5):
        k = int(n * frac)
        R0 = W.uniform_removals(n, k, seed=2031, device="cuda")
        R = torch.empty_like(R0)

        def restore():
            A.copy_(A0)
            R.copy_(R0)

        t_on = time_call(restore, lambda: ctx.partition(A, R), reps)
        t_rz = time_call(restore, lambda: ctx.remove(A, R), reps)
        emit({"config": "partition", "n": n, "frac": frac, "k": k, "partition_ms": round(t_on, 4),
              "redzone_ms": round(t_rz, 4),
              "partition_GBps": round((4 * n + 4 * (n - k)) / (t_on * 1e-3) / 1e9, 1)})
        del R0, R
        torch.cuda.empty_cache()


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--config", default="all", choices=["2", "3", "gather", "partition", "all"])
    p.add_argument("--reps", type=int, default=5)
    args = p.parse_args()
    import torch

    import redzone_amd as rz
    from redzone_amd import workloads as W

    ctx = rz.Context(0)

    def emit(d):
        print(json.dumps(d), flush=True)

    if args.config in ("2", "all"):
        sweep_config2(ctx, rz, W, torch, args.reps, emit)
    if args.config in ("3", "all"):
        sweep_config3(ctx, rz, W, torch, args.reps, emit)
    if args.config in ("gather", "all"):
        sweep_gather(ctx, rz, W, torch, args.reps, emit)
    if args.config in ("partition", "all"):
        sweep_partition(ctx, rz, W, torch, args.reps, emit)


if __name__ == "__main__":
    main()
