#!/usr/bin/env python3
"""Benchmark of Redzone O(k) in-place removal on one MI355X.

Workload (one step = one removal call on one list resident in HBM):
  n = 2^28 uint32 unique payloads, k = floor(n/100) uniform distinct removals
  (unsorted), seeded.  The call is destructive, so before every step the list
  and the removal list are restored from pristine copies and the caches are
  scrubbed by reading a 1 GiB buffer (cold-cache timing); the restore and the
  scrub are outside the timed call.

Printed: ONE JSON line on stdout.  `value` is list items per second of the
removal call (n / t).  `roofline` is for the dominant kernel of the launch
form the tuned table picks (rz_move_pair on this shape), timed live with the
library's events on the launch stream; `cpu_baseline` is the oracle's
restatement of Redzone_CPU timed on this host's cores (n/t, same workload).
Everything else goes to stderr.
"""
import argparse
import hashlib
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gitems/s & %HBM peak, n=2^28 k/n=1%; batched+learned-select vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--log2n", type=int, default=28)
    p.add_argument("--frac", type=float, default=0.01)
    p.add_argument("--pattern", default="uniform")
    p.add_argument("--elem-bytes", type=int, default=4)
    p.add_argument("--warm", action="store_true", help="skip the cache scrub before each call")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-extra", action="store_true", help="skip the comparator measurements")
    p.add_argument("--cpu-seconds", type=float, default=8.0, help="CPU time budget of each CPU-baseline measurement")
    p.add_argument("--variant", type=int, default=None,
                   help="0 three kernels, 1 cooperative, 2 persistent queue, 3 ordered, 4 fused, 5 dense")
    p.add_argument("--wg", type=int, default=256)
    p.add_argument("--items", type=int, default=None)
    p.add_argument("--max-slices", type=int, default=2048)
    p.add_argument("--slice-steps", type=int, default=2)
    p.add_argument("--no-batch", action="store_true", help="skip the batched (4096-list) comparison")
    p.add_argument("--mixed-lists", type=int, default=10000, help="lists of the mixed-stream evaluation (0 = skip)")
    p.add_argument("--chunk", type=int, default=1, help="with --variant: the config's chunk field (dense form: 2 = "
                   "bucketed tail marks)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run the counter calibration: copy, add to and sum a 1 GiB int32 tensor, three times")
    return p.parse_args()


def copy_calibration(torch):
    """Kernels of known traffic for scaling FETCH_SIZE / WRITE_SIZE: a 1 GiB blit
    copy (read 1 GiB, write 1 GiB), an in-place add (the same) and a sum."""
    n = 1 << 28
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    evict = torch.ones(n, dtype=torch.int32, device="cuda")
    for _ in range(3):
        evict.add_(1)
        b.copy_(a)
        evict.add_(1)
        b.add_(1)
        evict.add_(1)
        t = a.sum()
        torch.cuda.synchronize()
    log(f"calibration done ({int(t.item())})")


def tile_config(args, rz):
    """Explicit launch shape when --variant is given; otherwise the context's tuned table."""
    if args.variant is None:
        return None
    items = args.items or (8 if args.variant == 2 else 4)
    return rz.TileConfig(args.wg, items, args.max_slices, args.variant, args.slice_steps, args.chunk)


def algorithmic_bytes(R, n, k, s, torch):
    """B = 4k (read R) + s*k (read the tail) + s*(k - k_z) (write the moved items),
    and the move kernel's share when a separate kernel pairs the orphans."""
    sz = n - k
    tail = R >= sz
    kz = int(tail.sum().item())
    dead = torch.zeros(k, dtype=torch.bool, device=R.device)
    dead[(R[tail] - sz).long()] = True
    pairs = int(((~tail) & dead).sum().item())
    B = 4 * k + s * k + s * (k - kz)
    B_move = (4 + s) * k + s * (k - kz - pairs)
    return B, B_move, kz, pairs


# dominant kernel of each launch form on the headline shape, and whether it
# also moves the orphan pairs (then its algorithmic bytes are all of B); for
# the dense form the timed span is all of its kernels after the mark
# ("rz_dense_all")
KERNEL_OF_VARIANT = {0: ("rz_move", False), 4: ("rz_move_pair", True), 5: ("rz_dense_all", True)}
PATH_OF_VARIANT = {0: "rz_remove (mark, move, pair kernels)", 1: "rz_remove (one cooperative kernel)",
                   2: "rz_remove (persistent queue kernel)", 3: "rz_remove (one ordered launch)",
                   4: "rz_remove (mark kernel, then move with in-kernel orphan pairing)",
                   5: "rz_remove (dense form: bucketed marks and region-grouped writes)"}


def library_sha16():
    """First 16 hex digits of the SHA-256 of the loaded librz_amd.so (the build
    is deterministic, so the same sources give the same bytes)."""
    from redzone_amd import _native

    with open(_native.LIB_PATH, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def load_traffic(kernel, shape):
    """HBM bytes per launch of `kernel` from the committed counter record, or
    (None, reason) when the record was measured on another build or launch
    shape than the one this run loaded."""
    tpath = os.path.join(ROOT, "profiles", f"traffic_{kernel}_u32_n2e28_k1pct.json")
    if not os.path.exists(tpath):
        return None, "no counter record for this kernel"
    with open(tpath) as f:
        rec = json.load(f)
    sha = library_sha16()
    if rec.get("library_sha256_16") != sha:
        return None, f"counter record is of library {rec.get('library_sha256_16')}, this run loaded {sha}"
    if tuple(rec.get("launch_shape", ())) != tuple(shape):
        return None, f"counter record is of launch shape {rec.get('launch_shape')}, this run used {list(shape)}"
    return rec.get("hbm_bytes_per_launch"), None


def scatter_floor(torch, A, A0, R0, n, k, restore, scrub, steps, time_fn):
    """The same hole writes issued by a plain one-pass scatter (torch's
    index_copy_: read an index and a value, write the value at the index), with
    the holes in the removal list's order and sorted by address, cold like the
    timed calls.  Every correct removal writes these holes; their time is the
    floor of the scattered part of the call."""
    sz = n - k
    holes = R0[R0 < sz].long()
    vals = A0[sz: sz + holes.numel()].clone()
    ordered = torch.sort(holes).values
    t_rand, _, _ = time_fn(torch, None, restore, lambda: A.index_copy_(0, holes, vals), steps, 2, scrub, False)
    t_sort, _, _ = time_fn(torch, None, restore, lambda: A.index_copy_(0, ordered, vals), steps, 2, scrub, False)
    return {"holes": int(holes.numel()), "random_order_us": round(statistics.median(t_rand) * 1e3, 2),
            "address_order_us": round(statistics.median(t_sort) * 1e3, 2),
            "kind": "torch index_copy_ of the hole positions, cold cache, event-timed"}


def stage_timeline(ctx, restore, scrub, call, reps=5):
    """Mean stage timeline (us) of `reps` cold calls made with RZ_PROFILE
    (fused forms only; None otherwise)."""
    import redzone_amd as rz

    rows = []
    for _ in range(reps):
        restore()
        if scrub is not None:
            scrub()
        call(rz.RZ_PROFILE)
        try:
            rows.append(ctx.last_stages())
        except rz.RedzoneError:
            return None
    return {key: round(statistics.mean(r[key] for r in rows), 2) for key in rows[0]}


class Scrubber:
    def __init__(self, torch, nbytes=1 << 30):
        self.buf = torch.ones(nbytes // 8, dtype=torch.int64, device="cuda")

    def __call__(self):
        return self.buf.sum()


def time_calls(torch, ctx, fn_restore, fn_call, steps, warmup, scrub, profile):
    """Per-call event times (ms) and per-kernel profiles of `steps` calls."""
    times, profs = [], []
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for i in range(warmup + steps):
        fn_restore()
        if scrub is not None:
            scrub()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn_call()
        e1.record()
        if profile:
            prof = ctx.last_profile()
        e1.synchronize()
        if i >= warmup:
            times.append(e0.elapsed_time(e1))
            if profile:
                profs.append(prof)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    return times, profs, wall


def bench_batched(torch, ctx, rz, args):
    """Batched compaction: 4096 int32 lists, n log-uniform in [2^8, 2^20],
    k = floor(r n) with r ~ U[0, 0.3], uniform distinct removals.  One
    persistent launch (direct and hipGraph replay) vs 4096 plain per-list
    launches; inputs restored before each pass; wall time with syncs."""
    import numpy as np

    rng = np.random.default_rng(4096)
    specs = []
    for i in range(4096):
        n = int(2 ** rng.uniform(8, 20))
        specs.append((n, int(rng.uniform(0, 0.3) * n)))
    tot_n = sum(n for n, _ in specs)
    A0 = torch.empty(tot_n, dtype=torch.int32, device="cuda")
    ctx.gen_payload(A0, 77)
    A = torch.empty_like(A0)
    R0s, lists = [], []
    off = 0
    for i, (n, k) in enumerate(specs):
        R = torch.empty(k, dtype=torch.int32, device="cuda")
        ctx.gen_uniform_removals(R, n, 1000 + i)
        R0s.append(R)
        lists.append((A[off: off + n], torch.empty_like(R)))
        off += n
    sum_k = sum(k for _, k in specs)
    sum_B = 0
    for (n, k), R in zip(specs, R0s):
        if 0 < k < n:
            sum_B += 8 * k + 4 * (k - int((R >= n - k).sum().item()))

    def restore():
        A.copy_(A0)
        for (_, r), r0 in zip(lists, R0s):
            r.copy_(r0)
        torch.cuda.synchronize()

    def timed(fn, reps):
        ts = []
        for _ in range(reps + 1):
            restore()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts[1:]) * 1e3

    reps = max(3, args.steps // 4)
    b = ctx.batch(lists)
    info = b.info()
    t_batch = timed(b.launch, reps)
    s = torch.cuda.Stream()  # the legacy null stream cannot be captured
    with torch.cuda.stream(s):
        b.capture(stream=s)
    t_graph = timed(lambda: b.launch(stream=s), reps)
    b.close()
    t_plain = timed(lambda: [ctx.remove(a, r) for a, r in lists], reps)
    return {"lists": len(specs), "sum_n": tot_n, "sum_k": sum_k, "algorithmic_bytes": sum_B,
            "batched_ms": round(t_batch, 4), "graph_ms": round(t_graph, 4), "plain_per_list_ms": round(t_plain, 3),
            "batched_vs_plain": round(t_plain / min(t_batch, t_graph), 2),
            "batched_Gitems_s": round(tot_n / (min(t_batch, t_graph) * 1e-3) / 1e9, 1),
            "batched_GBps": round(sum_B / (min(t_batch, t_graph) * 1e-3) / 1e9, 1),
            "queue_items": info["nitems"], "grid": info["grid"], "config": ctx.batch_config().as_tuple()}


def bench_mixed(torch, ctx, args):
    """Mixed stream (lists of 4/8/16-byte items, n in [2^12, 2^26], k/n in
    [1e-5, 0.9], three removal patterns): the learned dispatcher (trained on a
    disjoint seed) vs the k < n/2 rule and vs the measured per-list best."""
    from redzone_amd import mixed as MX
    from redzone_amd.dispatch import MODEL_PATH, Dispatcher

    if not os.path.exists(MODEL_PATH):
        return None
    disp = Dispatcher.load(ctx, MODEL_PATH)
    specs = MX.stream_specs(args.mixed_lists, seed=2)
    import redzone_amd as rz

    untuned = rz.Context(ctx.device, load_tuned=False, load_dispatcher=False)
    res = MX.evaluate(ctx, disp, specs, 2 * 1000003, reps=3, log=log, ctx_untuned=untuned)
    res.pop("_per_list_us")
    res.pop("_chosen")
    out = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res.items()}
    out["lists"] = args.mixed_lists
    return out


def host_cpus():
    """The host's CPUs as this process sees them: the count the reference's
    hardware_concurrency() reports (all online CPUs), the affinity mask, the
    cgroup CPU quota and the model from lscpu."""
    info = {"os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        info["cgroup_cpus"] = None if quota == "max" else round(float(quota) / float(period), 2)
    except (OSError, ValueError):
        info["cgroup_cpus"] = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info["lscpu_" + key.strip().lower().replace("(s)", "s").replace(" ", "_")] = val.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(A0_h, R0_h, n, k, seconds):
    """The oracle's Redzone_CPU restatement on this host (O(k) restore between
    reps), with T = nproc - 1 threads (the reference's hardware_concurrency() - 1,
    Redzone_CPU.hxx:153) and with T = nproc; `value` is the T = nproc - 1 figure."""
    import numpy as np

    from oracle import pyoracle

    host = host_cpus()
    nproc = host["os_cpu_count"] or 1
    sz = n - k
    touched = R0_h[R0_h < sz].astype(np.int64)
    by_threads = {}
    A = None
    for threads in sorted({max(1, nproc - 1), nproc}):
        A = A0_h.copy()
        R = R0_h.copy()
        ts = []
        t_start = time.perf_counter()
        while len(ts) < 20 and (time.perf_counter() - t_start) < seconds:
            A[touched] = A0_h[touched]
            A[sz:] = A0_h[sz:]
            R[:] = R0_h
            st, t = pyoracle.redzone_cpu(A, R, threads=threads)
            assert st == 0
            ts.append(t[0] / 1e9)
        med = statistics.median(ts)
        by_threads[threads] = {"ms_median": round(med * 1e3, 3), "Gitems_s": round(n / med / 1e9, 3), "reps": len(ts)}
    t_ref = max(1, nproc - 1)
    ref = by_threads[t_ref]
    out = {"value": ref["Gitems_s"], "unit": "Gitems/s", "cores": t_ref, "kind": "port",
           "sample": f"full workload (n=2^{n.bit_length() - 1}, k={k}), {ref['reps']} reps, median "
                     f"{ref['ms_median']:.2f} ms, Redzone_CPU restatement (OpenMP, {t_ref} threads = nproc - 1 as "
                     f"Redzone_CPU.hxx:153)",
           "ms_median": ref["ms_median"], "by_threads": by_threads, "host": host}
    # the O(n) host comparator (std_parallel_remove.hxx:13-36 restated): full restore between reps
    B = A0_h.copy()
    ts2 = []
    t_start = time.perf_counter()
    while len(ts2) < 5 and (time.perf_counter() - t_start) < seconds:
        B[:] = A0_h
        st, t = pyoracle.remove_stable(B, R0_h, threads=t_ref)
        assert st == 0
        ts2.append(t / 1e9)
    med2 = statistics.median(ts2)
    out["std_parallel_remove"] = {"value": round(n / med2 / 1e9, 3), "unit": "Gitems/s", "cores": t_ref,
                                  "kind": "port", "ms_median": round(med2 * 1e3, 3),
                                  "sample": f"same workload, {len(ts2)} reps, mark + stable remove (OpenMP)"}
    return out, A


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(lists shard across GPUs without exchange; see the design notes). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if args.copy_calibration:
        copy_calibration(torch)
        return
    import redzone_amd as rz
    from redzone_amd import workloads as W

    ctx = rz.Context(0)
    cfg = tile_config(args, rz)
    info = ctx.device_info()
    n = 1 << args.log2n
    k = int(n * args.frac)
    s = args.elem_bytes
    used = cfg if cfg is not None else ctx.table(s, k.bit_length() - 1)
    kernel, kernel_pairs = KERNEL_OF_VARIANT.get(used.variant, ("rz_move", False))
    log(f"launch shape {used.as_tuple()} -> dominant kernel {kernel}")
    log(f"device {info}, n=2^{args.log2n}, k={k}, elem {s} B, pattern {args.pattern}")
    A0 = W.payload(s, n, seed=2026, device="cuda")
    R0 = W.removals(args.pattern, n, k, seed=2027, device="cuda")
    A = torch.empty_like(A0)
    R = torch.empty_like(R0)
    B, B_move, kz, pairs = algorithmic_bytes(R0, n, k, s, torch)
    log(f"k_z={kz}, orphan pairs={pairs}, algorithmic bytes {B / 1e6:.2f} MB (move kernel {B_move / 1e6:.2f} MB)")
    scrub = None if args.warm else Scrubber(torch)

    def restore():
        A.copy_(A0)
        R.copy_(R0)

    # timed calls: no profiling events inside the call
    times, _, wall = time_calls(torch, ctx, restore, lambda: ctx.remove(A, R, config=cfg), args.steps, args.warmup,
                                scrub, False)
    t_med = statistics.median(times)
    # per-kernel breakdown: the same calls with the library's events between its kernels
    _, profs, _ = time_calls(torch, ctx, restore, lambda: ctx.remove(A, R, flags=rz.RZ_PROFILE, config=cfg),
                             args.steps, 1, scrub, True)
    move_ms = statistics.mean(p[1] for p in profs)
    mark_ms = statistics.mean(p[0] for p in profs)
    pair_ms = statistics.mean(p[2] for p in profs)
    log(f"redzone: median {t_med * 1e3:.1f} us (p10 {sorted(times)[len(times) // 10] * 1e3:.1f}, "
        f"max {max(times) * 1e3:.1f}); kernels mark {mark_ms * 1e3:.1f} us, move {move_ms * 1e3:.1f} us, "
        f"pair {pair_ms * 1e3:.1f} us")

    # correctness of the last call against the oracle's canonical result (cheap: O(k) on the host)
    import numpy as np

    from oracle import pyoracle

    def _canonical(A_h, R_h):
        out = A_h.copy()
        assert pyoracle.canonical(out, R_h) == 0
        return out

    A0_h = A0.cpu().numpy()
    R0_h = R0.cpu().numpy()
    want = _canonical(A0_h, R0_h)
    got = A.cpu().numpy()
    bit_exact = bool(np.array_equal(got[: n - k], want[: n - k]))
    log(f"bit-exact against the oracle: {bit_exact}")
    try:  # every device-side check of the calls so far (orphan totals, bounded waits)
        device_errors = ctx.check_errors()
    except rz.RedzoneError as e:
        device_errors = e.error_flags

    B_kernel = B if kernel_pairs else B_move
    achieved = B_kernel / (move_ms * 1e-3) / 1e9
    op_gbs = B / (t_med * 1e-3) / 1e9
    traffic, traffic_note = None, "counters are recorded for the default workload only"
    if args.log2n == 28 and args.frac == 0.01 and s == 4 and args.pattern == "uniform" and cfg is None:
        traffic, traffic_note = load_traffic(kernel, used.as_tuple())
    # the floor of the scattered writes: the same holes written by a plain scatter
    floor = scatter_floor(torch, A, A0, R0, n, k, restore, scrub, max(5, args.steps // 2), time_calls)
    floor["kernel_us"] = round(move_ms * 1e3, 2)
    floor["frac_of_floor"] = round(floor["random_order_us"] / (move_ms * 1e3), 3)
    log(f"scatter floor: {floor}")
    stages = stage_timeline(ctx, restore, scrub, lambda f: ctx.remove(A, R, flags=f, config=cfg))
    log(f"stage timeline (us): {stages}")
    result = {
        "metric": METRIC,
        "value": round(n / (t_med * 1e-3) / 1e9, 2),
        "unit": "Gitems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_med, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {4: "u32", 8: "u64", 16: "rec16"}[s],
        "data": "synthetic (seeded unique payloads, uniform distinct removals)",
        "config": {"workload": f"redzone remove, n=2^{args.log2n} {s}-byte items, k/n={args.frac:g} {args.pattern}",
                   "n": n, "k": k, "k_z": kz, "orphan_pairs": pairs, "elem_bytes": s,
                   "cache": "warm" if args.warm else "cold (1 GiB read scrub before each call)",
                   "path": PATH_OF_VARIANT.get(used.variant, "rz_remove"), "launch_shape": used.as_tuple()},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": B_kernel, "avg_ms": round(move_ms, 5), "library": library_sha16(),
                     "scatter_floor": floor, **({"traffic_note": traffic_note} if traffic is None else {})},
        "op": {"algorithmic_bytes": B, "GBps": round(op_gbs, 1), "hbm_frac": round(op_gbs / HBM_PEAK_GBS, 4),
               "removals_per_s_G": round(k / (t_med * 1e-3) / 1e9, 3), "bit_exact": bit_exact,
               "device_error_bits": device_errors,
               "kernel_us": {"mark": round(mark_ms * 1e3, 2), "move": round(move_ms * 1e3, 2),
                             "pair": round(pair_ms * 1e3, 2)},
               "stages_us": stages,
               "wall_ms_per_step_incl_restore": round(wall * 1e3 / (args.steps + args.warmup), 3)},
    }

    if not args.no_extra:
        extra = {}
        # the O(n) stable partition on the same list
        t_on, _, _ = time_calls(torch, ctx, restore, lambda: ctx.partition(A, R), max(3, args.steps // 4), 1,
                                scrub, False)
        on_med = statistics.median(t_on)
        extra["on_partition"] = {"ms": round(on_med, 4), "Gitems_s": round(n / (on_med * 1e-3) / 1e9, 2),
                                 "redzone_speedup": round(on_med / t_med, 2),
                                 "bytes": 4 * k + s * n + s * (n - k) + n // 8}
        # k/n = 0.1%: Redzone vs the O(n) partition
        k2 = n // 1000
        R2 = W.removals(args.pattern, n, k2, seed=2028, device="cuda")
        R2w = torch.empty_like(R2)

        def restore2():
            A.copy_(A0)
            R2w.copy_(R2)

        t_rz2, _, _ = time_calls(torch, ctx, restore2, lambda: ctx.remove(A, R2w, config=cfg), args.steps, 2, scrub, False)
        t_on2, _, _ = time_calls(torch, ctx, restore2, lambda: ctx.partition(A, R2w), max(3, args.steps // 4), 1,
                                 scrub, False)
        rz2, on2 = statistics.median(t_rz2), statistics.median(t_on2)
        stages2 = stage_timeline(ctx, restore2, scrub, lambda f: ctx.remove(A, R2w, flags=f, config=cfg))
        # what bounds the ratio: one kernel launch of no work, and the bare scatter of this list's holes
        one = torch.zeros(1, dtype=torch.int32, device="cuda")
        t_launch, _, _ = time_calls(torch, ctx, lambda: None, lambda: one.add_(1), args.steps, 2, scrub, False)
        floor2 = scatter_floor(torch, A, A0, R2, n, k2, restore2, scrub, args.steps, time_calls)
        launch_us = statistics.median(t_launch) * 1e3
        # producer-fused marks: the tail bitmap is set before the clock starts (as a
        # producer of R that marks while it writes would leave it), and the call skips its mark step
        bits2 = torch.zeros((k2 + 31) // 32, dtype=torch.int32, device="cuda")

        def restore2_marked():
            restore2()
            ctx.mark_removals(R2w, n, bits2)

        t_pm, _, _ = time_calls(torch, ctx, restore2_marked, lambda: ctx.remove_premarked(A, R2w, bits2, config=cfg),
                                args.steps, 2, scrub, False)
        pm2 = statistics.median(t_pm)
        pm_exact = bool(np.array_equal(A.cpu().numpy()[: n - k2], _canonical(A0_h, R2.cpu().numpy())[: n - k2]))
        extra["k0.1pct"] = {"k": k2, "redzone_ms": round(rz2, 4), "on_partition_ms": round(on2, 4),
                            "redzone_vs_on": round(on2 / rz2, 1),
                            "premarked": {"ms": round(pm2, 4), "on_vs": round(on2 / pm2, 1), "bit_exact": pm_exact,
                                          "note": "tail marks set before the timed call, as by R's producer"},
                            "launch_shape": (cfg or ctx.table(s, k2.bit_length() - 1)).as_tuple(),
                            "stages_us": stages2, "one_kernel_launch_us": round(launch_us, 2),
                            "scatter_floor": floor2,
                            "on_vs_scatter_floor": round(on2 * 1e3 / floor2["random_order_us"], 1)}
        # warm-cache figure of the headline call
        t_warm, _, _ = time_calls(torch, ctx, restore, lambda: ctx.remove(A, R, config=cfg), args.steps, 2, None, False)
        extra["warm_ms"] = round(statistics.median(t_warm), 4)
        # the same call with n and k read on the device (rz_remove_sized, one-launch fused form)
        d_n = torch.tensor([n], dtype=torch.int32, device="cuda")
        d_k = torch.tensor([k], dtype=torch.int32, device="cuda")
        t_sz, _, _ = time_calls(torch, ctx, restore, lambda: ctx.remove_sized(A, R, d_n, d_k), args.steps, 2, scrub,
                                False)
        sz_med = statistics.median(t_sz)
        extra["device_sized"] = {"ms": round(sz_med, 4), "vs_remove": round(sz_med / t_med, 3),
                                 "bit_exact": bool(np.array_equal(A.cpu().numpy()[: n - k], want[: n - k]))}
        result["comparators"] = extra
        log(f"comparators: {extra}")

    if not args.no_batch:
        del A, R
        torch.cuda.empty_cache()
        result["batched"] = bench_batched(torch, ctx, rz, args)
        log(f"batched: {result['batched']}")
    if args.mixed_lists > 0:
        result["mixed"] = bench_mixed(torch, ctx, args)
        log(f"mixed: {result['mixed']}")

    if not args.no_cpu:
        # restore the pristine list on the host copy for the CPU baseline
        cb, A_cpu = cpu_baseline(A0_h, R0_h, n, k, args.cpu_seconds)
        cb["matches_gpu"] = bool(np.array_equal(A_cpu[: n - k], want[: n - k]))
        result["cpu_baseline"] = cb
        log(f"cpu baseline: {cb}")
    else:
        result["cpu_baseline"] = None

    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
