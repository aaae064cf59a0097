#!/usr/bin/env python3
"""Autotunes the launch shapes on this device and writes the table that
Context loads by default (redzone_amd/tuned_gfx950.txt).

    python -m redzone_amd.tune [--out PATH] [--reps 5]
    python -m redzone_amd.tune --check-list-bytes 16777216 67108864

For every element size (4, 8, 16 bytes) and every k bucket [2^b, 2^(b+1))
(up to k = n/2) the native tuner (rz_tune) times the built-in candidate
shapes at k = 2^(b+1/2) on a synthetic list of 1 GiB with cold caches and
keeps the fastest; then the batch tuner
(rz_tune_batch) picks the queue shape on a 4096-list batch shaped like the
batched benchmark.  Every line of the result is printed as JSON.

The table is keyed by element size and log2 k only.  --check-list-bytes
measures what that costs on lists of other sizes: for each given list size
(bytes) and each bucket it times every candidate on a cold list of that size
and prints the installed shape's time next to the fastest candidate's, then
the geometric mean of their ratio.  It writes no table.
"""
import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def check_sizes(rz, sizes, reps):
    import math

    installed = rz.Context(0)
    scratch = rz.Context(0, load_tuned=False)  # rz_tune installs its winner here, not in `installed`
    for list_bytes in sizes:
        ratios = []
        for esz in (4, 8, 16):
            n = list_bytes // esz
            for b in range(8, 32):
                k = int(2 ** (b + 0.5))
                if k * 2 > n:
                    break
                entry = installed.table(esz, b).as_tuple()
                best, ms, results = scratch.tune(esz, n, k, reps=reps, cold=True)
                t = [tm for c, tm in results if c.as_tuple() == entry and tm > 0]
                row = {"list_bytes": list_bytes, "elem_bytes": esz, "n": n, "log2k": b, "installed": entry,
                       "installed_ms": round(t[0], 5) if t else None, "best": best.as_tuple(), "best_ms": round(ms, 5)}
                if t:
                    ratios.append(t[0] / ms)
                    row["ratio"] = round(t[0] / ms, 3)
                print(json.dumps(row), flush=True)
        if ratios:
            print(json.dumps({"list_bytes": list_bytes, "buckets": len(ratios),
                              "geomean_ratio": round(math.exp(sum(map(math.log, ratios)) / len(ratios)), 4),
                              "max_ratio": round(max(ratios), 3)}), flush=True)


def main():
    import redzone_amd as rz
    from redzone_amd import _native

    p = argparse.ArgumentParser()
    p.add_argument("--out", default=_native.TUNED_TABLE)
    p.add_argument("--reps", type=int, default=5)
    p.add_argument("--check-list-bytes", type=int, nargs="*", default=None)
    args = p.parse_args()
    if args.check_list_bytes:
        check_sizes(rz, args.check_list_bytes, args.reps)
        return
    ctx = rz.Context(0, load_tuned=False)
    print(json.dumps({"device": ctx.device_info()}), flush=True)
    lo_b = 8
    for esz in (4, 8, 16):
        n = (1 << 30) // esz
        hi_b = lo_b
        for b in range(lo_b, 32):
            k = int(2 ** (b + 0.5))  # the middle of bucket [2^b, 2^(b+1)) on a log scale
            if k * 2 > n:
                break
            best, ms, results = ctx.tune(esz, n, k, reps=args.reps, cold=True)
            hi_b = b
            print(json.dumps({"elem_bytes": esz, "n": n, "log2k": b, "best": best.as_tuple(), "best_ms": round(ms, 5),
                              "all": [[c.as_tuple(), round(t, 5)] for c, t in results]}), flush=True)
        # buckets outside the tuned range take the nearest tuned shape
        for bb in range(hi_b + 1, 33):
            ctx.set_table(esz, bb, ctx.table(esz, hi_b))
        for bb in range(0, lo_b):
            ctx.set_table(esz, bb, ctx.table(esz, lo_b))
    best, ms, results = ctx.tune_batch(4, 4096, 1 << 8, 1 << 20, 0.3, reps=max(2, args.reps // 2))
    print(json.dumps({"batch": best.as_tuple(), "best_ms": round(ms, 4),
                      "all": [[c.as_tuple(), round(t, 4)] for c, t in results]}), flush=True)
    ctx.save_table(args.out)
    print(json.dumps({"saved": args.out}), flush=True)


if __name__ == "__main__":
    main()
