#!/usr/bin/env python3
"""Benchmark: M2L translations/s at P=10, FP64, on one MI355X.

Default workload: a batch of 100,000 independent M2L translations of order
P=10 (degrees 0..9, 110 reals per expansion), each with its own source and
target.  Sources: Re/Im i.i.d. U[-1,1] with Im C_n^0 = 0; shifts: uniform
directions with |x| ~ U[2,4] (M2L) or the octant offsets (+-0.5)^3 scaled by
U[0.5,1] (M2M, L2L); targets accumulate.  Inputs are seeded
(solidfmm_amd.workload) and resident in HBM before the timed region.

A step is one pass of the plan over the batch (one kernel launch).  W warmup
steps, then K timed steps between device synchronisations; HIP events on the
launch stream give the per-step intervals, the isolated batch latency and
the kernel's average duration for the roofline.

--op / --P / --n / --kernels select other workloads (e.g. --op m2m --P 16
--n 1000000 for the M2M/L2L configuration); the default line is the headline.
Prints one JSON line.  Only one GPU is supported: --gpus N > 1 exits.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

OPS = {"m2m": 0, "m2l": 1, "l2l": 2}
KERNELS = {"auto": 0, "lane": 1, "mfma": 2}


def algorithmic_bytes(p_in, p_out):
    """Read the source once, read and write the target, plus the 3 FP64
    components of the shift: 8*[P_in(P_in+1) + 2 P_out(P_out+1)] + 24."""
    return 8 * (p_in * (p_in + 1) + 2 * p_out * (p_out + 1)) + 24


def algorithmic_flops(op, P):
    """4 S2 per side for the swaps ((n+1)^2 FMA per swap, 4 swaps per
    degree), 14 T per side for the rotations, plus the z-translation:
    Hankel (M2L) or triangular Toeplitz (M2M, L2L) sums."""
    s2 = sum(k * k for k in range(1, P + 1))
    t = P * (P + 1) // 2
    if op == "m2l":
        fz = sum(4 * (P - m) * (P - m) for m in range(P))
    else:
        fz = sum(2 * (P - m) * (P - m + 1) for m in range(P))
    return 4 * (s2 + s2) + 14 * (t + t) + fz


# Host threads of the CPU baseline: one per physical core of the affinity
# mask, at most the CPU share a one-GPU job gets on the GPU hosts (16 CPUs),
# unless --cpu-threads says otherwise.
HOST_SHARE = 16


def reference_pool(ref, op, P, items, src, dst, threads, repeats=5):
    """Median wall time of the compiled reference's batch protocol (pinned
    worker pool, warm-up on the first 1024 items, `repeats` timed calls over
    the list; oracle/ref_shim.cpp) and the host description for the line."""
    info = ref.host_info()
    if not threads:
        threads = min(info["physical_cores"], HOST_SHARE)
    secs, used, pinned = ref.pool_translate(op, P, items, src, dst, threads=threads, repeats=repeats)
    info.update({"threads": used, "threads_pinned": pinned,
                 "batch_latency_ms": {"p50": float(np.median(secs)) * 1e3, "max": float(np.max(secs)) * 1e3,
                                      "min": float(np.min(secs)) * 1e3, "calls": len(secs)}})
    return float(np.median(secs)), info


def cpu_baseline(op, items, src, gpu_out, P, threads):
    """Time the reference's CPU path (oracle/_ref) on the same batch; fall back
    to the oracle port when the reference library was not built."""
    import oracle
    n = len(items)
    opi = OPS[op]
    dst = np.zeros(n * P * (P + 1))
    if oracle.Reference.available():
        ref = oracle.Reference()
        secs, info = reference_pool(ref, opi, P, items, src, dst, threads)
        # one thread on a slice of the batch: the per-core rate
        n1 = min(n, 20000)
        secs1, _ = reference_pool(ref, opi, P, items[:n1], src, np.zeros(n1 * P * (P + 1)), 1, repeats=3)
        kind, cores = "reference", info["threads"]
        note = (f"vorticle/solidfmm 1.5, {op}_unchecked on a pool of {cores} worker threads pinned one per "
                "physical core (the protocol of multithreaded_translator, which needs hwloc, absent here), "
                "targets split by hash range, 1337-item chunks; warm-up on 1024 items, then the whole "
                "batch, median of 5")
        extra = dict(info, median_s=secs, per_thread=n1 / secs1,
                     projected_all_physical_cores=n1 / secs1 * info["physical_cores"])
        # The reference's AVX-512 lanes 12-15 defect (microkernel_avx512f.cpp:2182)
        # makes its whole-batch output unfit for comparison; check the GPU
        # against the reference in calls of <= 12 items instead.
        nchk = min(n, 4096)
        check = np.zeros(nchk * P * (P + 1))
        sub = items[:nchk].copy()
        sub["dst"] = np.arange(nchk, dtype=np.uint64) * np.uint64(P * (P + 1))
        ref.translate(opi, P, sub, src, check, threads=1, repeats=1, chunk=ref.SAFE_CHUNK)
    else:
        orc = oracle.Oracle()
        t0 = time.perf_counter()
        orc.translate(opi, items, src, dst)
        secs = time.perf_counter() - t0
        kind, cores, note, extra = "port", 1, "oracle FP64 restatement, 1 thread, whole batch", {}
        nchk = min(n, 4096)
        check = dst[: nchk * P * (P + 1)]
    got = gpu_out.reshape(n, -1)[:nchk]
    err = oracle.degree_errors(got, check.reshape(nchk, -1), P)
    return dict({"value": n / secs, "unit": "translations/s", "cores": cores, "kind": kind, "sample": note},
                **extra), err


def truth_errors(op, items, src, gpu_out, P, nchk=4096):
    """Per-item errors (max over degrees of the per-degree normwise relative
    error) of the GPU, the compiled reference and the FP64 oracle against
    the binary128 truth (oracle/truth.c) on the batch's first nchk items."""
    import oracle
    opi = OPS[op]
    nchk = min(nchk, len(items))
    sub = items[:nchk].copy()
    sub["dst"] = np.arange(nchk, dtype=np.uint64) * np.uint64(P * (P + 1))
    dst = np.zeros(nchk * P * (P + 1))
    hi, lo = oracle.Truth().translate(opi, sub, src, len(dst))
    outs = {"gpu": gpu_out.reshape(len(items), -1)[:nchk].reshape(-1),
            "oracle": oracle.Oracle().translate(opi, sub, src, np.zeros_like(dst))}
    if oracle.Reference.available():
        r = np.zeros_like(dst)
        ref = oracle.Reference()
        ref.translate(opi, P, sub, src, r, threads=1, repeats=1, chunk=ref.SAFE_CHUNK)
        outs["reference"] = r
    rep = {"items": nchk}
    for k, v in outs.items():
        e = oracle.truth_degree_errors(v, hi, lo, P).max(axis=1)
        rep[k] = {"p50": float(np.percentile(e, 50)), "p99": float(np.percentile(e, 99)), "max": float(e.max())}
    return rep


def arith_mode(args):
    """Issue rates of the arithmetic the precision modes trade against
    (sfmm_probe_arith), and what they imply for a double-word FP32 path."""
    import ctypes
    import solidfmm_amd
    lib = solidfmm_amd.load()
    names = ["fp64_fma_valu", "fp32_fma_valu", "dw32_fma_valu", "fp64_mfma_16x16x4", "fp32_mfma_16x16x4"]
    rates = {}
    for k, name in enumerate(names):
        v = ctypes.c_double()
        rc = lib.sfmm_probe_arith(0, k, ctypes.byref(v))
        if rc != 0:
            raise RuntimeError(f"sfmm_probe_arith({k}) failed ({rc})")
        rates[name] = v.value
    line = {"mode": "arith", "unit": "flop/s", "rates": rates,
            "ratios": {"fp32_valu_over_fp64_valu": rates["fp32_fma_valu"] / rates["fp64_fma_valu"],
                       "dw32_valu_over_fp64_valu": rates["dw32_fma_valu"] / rates["fp64_fma_valu"],
                       "fp32_mfma_over_fp64_mfma": rates["fp32_mfma_16x16x4"] / rates["fp64_mfma_16x16x4"],
                       # hi/lo split products need 3 FP32 products per FP64 one
                       # (and still accumulate in FP32); exact 12-bit slices of
                       # 48-bit operands need 10
                       "split3_mfma_over_fp64_mfma": rates["fp32_mfma_16x16x4"] / 3 / rates["fp64_mfma_16x16x4"],
                       "slices10_mfma_over_fp64_mfma": rates["fp32_mfma_16x16x4"] / 10 / rates["fp64_mfma_16x16x4"]}}
    print(json.dumps(line), flush=True)


def copy_mode(args):
    """Plain device-to-device copy of known size: reads and writes
    args.copy_mib MiB per launch."""
    import torch
    n = args.copy_mib * 2 ** 20 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    reps = max(1, args.steps)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 2 * n * 8
    print(json.dumps({"mode": "copy", "bytes_per_launch": nbytes, "read_bytes": n * 8, "write_bytes": n * 8,
                      "ms_per_launch": ms, "GBps": nbytes / (ms * 1e-3) / 1e9}), flush=True)


def mixed_mode(args):
    """Mixed-order M2L: (P_in, P_out) i.i.d. uniform on {pmin..pmax}, one
    target each.  The autotuner picks a kernel family per order bucket
    max(P_in, P_out); the plan (one launch per bucket) is captured once into
    a HIP graph and each step is one graph replay."""
    import torch

    import solidfmm_amd
    from solidfmm_amd import workload as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    solidfmm_amd.load()
    n = args.n if args.n != 100_000 else 500_000
    seed = 0x5011DF02
    items, src, dst = W.mixed_batch(n, seed, args.pmin, args.pmax)
    ctx = solidfmm_amd.Context(args.pmax)
    ctx.set_kernels(KERNELS[args.kernels])
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    if not args.no_autotune and args.kernels == "auto":
        ctx.autotune(W.M2L, items, stream)
    tune_s = time.perf_counter() - t0
    plan = solidfmm_amd.Plan(ctx, W.M2L, items)
    s = torch.from_numpy(src).cuda()
    d = torch.zeros(len(dst), dtype=torch.float64, device="cuda")
    plan.execute(s, d, stream)
    torch.cuda.synchronize()
    gpu_out = d.cpu().numpy()
    if args.no_graph:
        step = lambda: plan.execute(s, d, stream)  # noqa: E731
    else:
        plan.capture(s, d)
        step = lambda: plan.replay(stream)  # noqa: E731
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    K = args.steps
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for k in range(K):
        step()
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = ev[0].elapsed_time(ev[K])
    lat = []
    for _ in range(args.latency_reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        step()
        b.record(stream)
        b.synchronize()
        lat.append(a.elapsed_time(b))
    lat = np.array(lat) if lat else np.zeros(1)
    ms = total_ms / K
    value = n / (ms * 1e-3)
    pin = items["p_in"].astype(np.int64)
    pout = items["p_out"].astype(np.int64)
    nbytes = float(np.sum(8 * (pin * (pin + 1) + 2 * pout * (pout + 1)) + 24))
    achieved = nbytes / (ms * 1e-3) / 1e9
    launches = plan.launch_info()
    line = {
        "metric": f"mixed-order M2L translations/s (orders {args.pmin}..{args.pmax}) FP64 on 1 GPU",
        "value": value, "unit": "translations/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"M2L FP64, {n} translations, P_in and P_out i.i.d. uniform on "
                               f"{{{args.pmin}..{args.pmax}}}, |shift| in [2,4], one target each, seeded; "
                               + ("one HIP graph replay per step" if not args.no_graph else "plan execute per step"),
                   "batch": n, "pmin": args.pmin, "pmax": args.pmax, "kernels": args.kernels,
                   "autotuned": not args.no_autotune and args.kernels == "auto", "autotune_s": tune_s,
                   "launches": [{"family": {1: "lane", 2: "mfma", 3: "generic"}[f], "order": o, "per_item_orders": r,
                                 "items": c} for f, o, r, c in launches]},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "wall_s": wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                     "traffic": None, "algorithmic_bytes_per_step": nbytes,
                     "note": "all launches of the step (one per order bucket) together"},
    }
    if not args.no_cpu_baseline:
        import oracle
        if oracle.Reference.available():
            ref = oracle.Reference()
            out = np.zeros_like(dst)
            secs, info = reference_pool(ref, W.M2L, args.pmax, items, src, out, args.cpu_threads, repeats=3)
            line["cpu_baseline"] = dict({"value": n / secs, "unit": "translations/s", "cores": info["threads"],
                                         "kind": "reference",
                                         "sample": f"vorticle/solidfmm 1.5 m2l_unchecked on {info['threads']} pinned "
                                                   "worker threads (multithreaded_translator's protocol), warm-up "
                                                   "1024 items, whole batch, median of 3"}, **info)
        nchk = 2000
        idx = np.linspace(0, n - 1, nchk).astype(np.int64)
        sub = items[idx]
        want = oracle.Oracle().translate(W.M2L, sub, src, np.zeros_like(dst))
        e = oracle.mixed_errors(gpu_out, want, sub)
        line["rel_err_vs_oracle"] = {"p50": float(np.percentile(e, 50)), "p99": float(np.percentile(e, 99)),
                                     "max": float(e.max()), "sample": nchk}
    print(json.dumps(line), flush=True)


def precision_mode(args):
    """Precision study: M2L at P in --orders (default 10, 16, 24), --n
    translations each (default 200,000), in FP64, FP32 and split FP32
    (double-word).  For each: translations/s over K timed steps and the
    per-translation error (max over degrees of the per-degree normwise
    relative error) against the FP64 oracle on a 2,000-item sample."""
    import torch

    import oracle
    import solidfmm_amd
    from solidfmm_amd import workload as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    solidfmm_amd.load()
    n = args.n if args.n != 100_000 else 200_000
    orders = [int(p) for p in args.orders.split(",")]
    stream = torch.cuda.current_stream()
    orc = oracle.Oracle()
    results = []
    for P in orders:
        items, src, dst = W.uniform_batch(W.M2L, P, P, n, 0x5011DF03 + P)
        s = torch.from_numpy(src).cuda()
        idx = np.linspace(0, n - 1, 2000).astype(np.int64)
        sub = items[idx].copy()
        sub["dst"] = np.arange(len(idx), dtype=np.uint64) * np.uint64(P * (P + 1))
        want = orc.translate(W.M2L, sub, src, np.zeros(len(idx) * P * (P + 1)))
        naive = orc.translate(W.M2L, sub, src, np.zeros(len(idx) * P * (P + 1)), naive=True)
        spread = oracle.degree_errors(naive, want, P).max(axis=1)
        for name, precision in solidfmm_amd.PRECISIONS.items():
            ctx = solidfmm_amd.Context(P, precision=precision)
            plan = solidfmm_amd.Plan(ctx, W.M2L, items)
            d = torch.zeros(n * P * (P + 1), dtype=torch.float64, device="cuda")
            plan.execute(s, d, stream)
            torch.cuda.synchronize()
            got = d.cpu().numpy().reshape(n, -1)[idx]
            e = oracle.degree_errors(got, want.reshape(len(idx), -1), P).max(axis=1)
            for _ in range(args.warmup):
                plan.execute(s, d, stream)
            K = args.steps
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(K):
                plan.execute(s, d, stream)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / K
            fam = plan.launch_info()[0][0]
            results.append({"P": P, "precision": name,
                            "kernel": {1: "lane", 2: "mfma", 3: "generic"}[fam],
                            "translations_per_s": n / (ms * 1e-3), "ms_per_step": ms,
                            "hbm_frac": n / (ms * 1e-3) * algorithmic_bytes(P, P) / 8e12,
                            "rel_err_vs_fp64_oracle": {"p50": float(np.percentile(e, 50)),
                                                       "p99": float(np.percentile(e, 99)),
                                                       "max": float(e.max())},
                            "oracle_vs_naive_p99": float(np.percentile(spread, 99))})
            print(json.dumps(results[-1]), file=sys.stderr, flush=True)
            del plan, ctx, d
    head = [r for r in results if r["P"] == orders[0] and r["precision"] == "split32"]
    line = {"metric": "M2L precision study: translations/s and error vs the FP64 oracle, FP64 / FP32 / split FP32",
            "value": head[0]["translations_per_s"] if head else None, "unit": "translations/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64/f32/f32x2", "data": "synthetic",
            "config": {"workload": f"M2L, P in {orders}, {n} translations each, |shift| in [2,4], seeded; FP64 "
                                   "pools, geometry and accumulation in every mode", "batch": n, "orders": orders},
            "results": results}
    print(json.dumps(line), flush=True)


def serving_mode(args):
    """Serving: the FMM passes (M2M up, M2L, L2L down) of an adaptive octree
    over a Plummer sphere (--points, at most 64 per leaf; orders 20 at the
    root to 6 at the leaves) cut into requests of U[1k, 50k] translations.
    The kernel families are autotuned on the workload's own lists (FP64).
    One-at-a-time GPU time is measured first (each request alone,
    synchronised); then, after one untimed pass, saturated runs of the
    continuous-batching loop
    (--slots streams, --budget-mib device arena; every request available at
    once; the median of three) give the loop's capacity, and the requests
    arrive as a Poisson process at --load x that capacity.  Reports translations/s and p50/p99
    request latency per precision."""
    import torch

    import solidfmm_amd
    from solidfmm_amd import octree as O
    from solidfmm_amd import serving as S
    from solidfmm_amd import workload as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    solidfmm_amd.load()
    seed = 0x5011DF04
    t0 = time.perf_counter()
    tree = O.Octree(O.plummer(args.points, seed), leaf_size=64)
    if args.lists == "mac":
        from solidfmm_amd import fmm as F
        P, off, pool_len, passes = F.mac_translation_lists(tree, O.plummer(args.points, seed), args.theta, 20, 6)
    else:
        P, off, pool_len, passes = O.translation_lists(tree, 20, 6)
    m0 = W.uniform(seed, 7, pool_len, -1.0, 1.0)
    build_s = time.perf_counter() - t0
    results = []
    for name in args.precisions.split(","):
        precision = solidfmm_amd.PRECISIONS[name]
        ctx = solidfmm_amd.Context(20, precision=precision)
        if not args.no_autotune and precision == solidfmm_amd.SFMM_FP64:
            # kernel family per order bucket, timed on the workload's own lists
            for op in (W.M2M, W.M2L, W.L2L):
                its = [p[1] for p in passes if p[0] == op and len(p[1])]
                if its:
                    ctx.autotune(op, np.concatenate(its))
        pools = (torch.from_numpy(m0).cuda(), torch.zeros(pool_len, dtype=torch.float64, device="cuda"))
        reqs = S.make_requests(passes, seed)
        S.run_sequential(ctx, pools, reqs[:20])  # warm-up
        gpu_ms = S.run_sequential(ctx, pools, reqs)
        total = sum(len(r.items) for r in reqs)
        cap_rps = len(reqs) / (sum(gpu_ms) * 1e-3)
        row = {"precision": name, "requests": len(reqs), "translations": total,
               "alone_gpu_ms": sum(gpu_ms), "alone_translations_per_s": total / (sum(gpu_ms) * 1e-3)}
        cls = S.NativeServer if args.server == "native" else S.Server
        server = cls(ctx, pools, len(tree), slots=args.slots, budget_bytes=args.budget_mib << 20)
        # one untimed pass warms the loop (first use of its arena, pinned
        # pool, threads and the plan caches)
        server.run(S.make_requests(passes, seed), realtime=False)
        # the loop's own capacity: requests/s with every request available
        # (dependencies between levels keep it below the one-at-a-time sum),
        # the median of three saturated runs (single runs vary by about 15 %,
        # and a high one sets an offered rate the loop cannot sustain)
        sat_runs = []
        for _ in range(3):
            reqs = S.make_requests(passes, seed)
            sat_runs.append(server.run(reqs, realtime=False))
        sat_runs.sort(key=lambda r: r["span_ms"])
        row["saturated"] = sat_runs[1]
        row["saturated_spans_ms"] = [r["span_ms"] for r in sat_runs]
        sat_rps = len(reqs) / (row["saturated"]["span_ms"] * 1e-3)
        reqs = S.make_requests(passes, seed)
        S.poisson_arrivals(reqs, args.load * sat_rps, seed)
        row["rate_requests_per_s"] = args.load * sat_rps
        row["poisson"] = server.run(reqs, realtime=True)
        results.append(row)
        print(json.dumps(row), file=sys.stderr, flush=True)
        del server, pools, ctx
    head = results[0]["poisson"]
    line = {"metric": "serving: translations/s and p50/p99 request latency, FMM passes of a Plummer octree",
            "value": head["translations_per_s"], "unit": "translations/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 / f32x2",
            "data": "synthetic",
            "config": {"workload": f"Plummer sphere of {args.points} points (a=1, seeded), octree <= 64 per leaf "
                                   f"({len(tree)} cells, depth {tree.depth}), orders 20 (root) to 6 (leaves); "
                                   + (f"M2L lists by dual tree walk, theta={args.theta}, smallest-ball centres; "
                                      if args.lists == "mac" else "M2L by same-level V-lists; ")
                                   + "M2M/M2L/L2L passes cut into requests of U[1k,50k] translations, "
                                   + f"{args.server} request loop, Poisson "
                                   f"arrivals at {args.load} x the loop's saturated capacity; {args.slots} streams, "
                                   f"{args.budget_mib} MiB device arena for requests in flight",
                       "points": args.points, "cells": int(len(tree)), "depth": int(tree.depth),
                       "tree_and_lists_s": build_s},
            "latency_ms": head["latency_ms"], "results": results}
    print(json.dumps(line), flush=True)


def fmm_mode(args):
    """Whole far-field FMM evaluation of a Plummer sphere (--points, leaf
    size 64, acceptance parameter --theta, orders 20 at the root to 6 at the
    leaves): P2M, M2M, M2L, L2L and L2P on the GPU, timed over --steps
    evaluations; the error of potential and gradient (far field plus the
    host near field) against direct sums on 64 sampled points."""
    import torch

    import solidfmm_amd
    from solidfmm_amd import octree as O
    from solidfmm_amd.fmm import FMM

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    solidfmm_amd.load()
    seed = 0x5011DF04
    rng = np.random.default_rng(seed)
    pts = np.c_[O.plummer(args.points, seed), rng.uniform(0.5, 1.5, args.points) / args.points]
    t0 = time.perf_counter()
    ctx = solidfmm_amd.Context(20)
    fmm = FMM(ctx, pts, leaf_size=64, theta=args.theta, p_root=20, p_leaf=6, autotune=not args.no_autotune)
    setup_s = time.perf_counter() - t0
    for _ in range(args.warmup):
        fmm.evaluate()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        fmm.evaluate()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    marks = []

    def mark(name):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        marks.append((name, ev))

    phi, grad = fmm.evaluate(mark=mark)
    torch.cuda.synchronize()
    phases = {}
    for (_, a), (name, b) in zip(marks[:-1], marks[1:]):
        phases[name] = phases.get(name, 0.0) + a.elapsed_time(b)
    phi = phi.cpu().numpy()
    grad = grad.cpu().numpy()
    ep, eg = [], []
    for i in rng.choice(args.points, 64, replace=False):
        d = pts[i, :3] - pts[:, :3]
        r = np.linalg.norm(d, axis=1)
        r[i] = np.inf
        want_p = (pts[:, 3] / r).sum()
        want_g = -(pts[:, 3, None] * d / r[:, None] ** 3).sum(0)
        near_p, near_g = fmm.near_field_at(pts, i)
        ep.append(abs(phi[i] + near_p - want_p) / abs(want_p))
        eg.append(np.linalg.norm(grad[i] + near_g - want_g) / np.linalg.norm(want_g))
    n_tr = fmm.translations
    cpu = None
    if not args.no_cpu_baseline:
        # the reference's translations of one evaluation (M2M, M2L and L2L
        # passes in order) on the host cores; random pools of the same size
        import oracle
        if oracle.Reference.available():
            ref = oracle.Reference()
            mpool = rng.uniform(-1, 1, fmm.pool_len)
            lpool = np.zeros(fmm.pool_len)
            secs = 0.0
            for op, it, _, _, _ in fmm.passes:
                src = lpool if op == 2 else mpool
                dst = mpool if op == 0 else lpool
                t, info = reference_pool(ref, op, 20, it, src, dst, args.cpu_threads, repeats=1)
                secs += t
            cpu = dict({"value": n_tr / secs, "unit": "translations/s", "cores": info["threads"], "kind": "reference",
                        "sample": f"vorticle/solidfmm 1.5 *_unchecked on {info['threads']} pinned worker threads "
                                  "(multithreaded_translator's protocol), the M2M/M2L/L2L passes of one evaluation "
                                  f"in order (P2M/L2P not timed); {secs:.2f} s"}, **info)
    line = {"metric": "far-field FMM evaluations of a Plummer sphere: translations/s over P2M..L2P",
            "value": n_tr / (ms * 1e-3), "unit": "translations/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"Plummer sphere of {args.points} points (a=1, seeded, charges U[0.5,1.5]/n), "
                                   f"octree <= 64 per leaf ({len(fmm.tree)} cells, depth {fmm.tree.depth}), "
                                   f"smallest-ball centres, dual tree walk theta={args.theta}, orders 20 (root) to "
                                   "6 (leaves); one step = P2M + M2M + M2L + L2L + L2P (near field not included)",
                       "points": args.points, "cells": int(len(fmm.tree)), "translations": n_tr,
                       "m2l": int(len(fmm.m2l_src)), "near_pairs": int(len(fmm.near_src)), "setup_s": setup_s},
            "points_per_s": args.points / (ms * 1e-3), "phase_ms": phases, "cpu_baseline": cpu,
            "rel_err_vs_direct": {"potential_max": float(np.max(ep)), "potential_p50": float(np.median(ep)),
                                  "gradient_max": float(np.max(eg)), "gradient_p50": float(np.median(eg))}}
    print(json.dumps(line), flush=True)


def host_mode(args):
    """The host-memory path (PCIe-inclusive): expansions live in host pools
    and every call copies the distinct source and target blocks in, runs the
    plan and copies the targets back (sfmm_staging_* / sfmm_translate_host,
    the path of the C++ drop-in).  Two lists: the headline batch (--n
    independent M2L at --P, every block moved once per translation) and the
    M2L list of a Plummer FMM (--points, theta --theta, orders 20..6), where
    sources are shared by many targets.  Reports translations/s of
    (a) pinned staging pools already packed, (b) pageable host pools
    (translate_host, which packs them with host threads)."""
    import torch

    import solidfmm_amd
    from solidfmm_amd import fmm as F
    from solidfmm_amd import octree as O
    from solidfmm_amd import workload as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    solidfmm_amd.load()
    rows = []
    items, src, dst = W.uniform_batch(W.M2L, args.P, args.P, args.n, 0x5011DF00)
    lists = [("independent M2L P=%d, %d items" % (args.P, args.n), args.P, items, src, dst)]
    tree = O.Octree(O.plummer(args.points, 0x5011DF04), leaf_size=64)
    P, off, pool_len, passes = F.mac_translation_lists(tree, O.plummer(args.points, 0x5011DF04), args.theta, 20, 6)
    m2l = [p for p in passes if p[0] == W.M2L][0][1]
    m0 = W.uniform(0x5011DF04, 7, pool_len, -1.0, 1.0)
    lists.append(("Plummer FMM M2L list, %d points, theta %.2f, orders 20..6" % (args.points, args.theta), 20, m2l,
                  m0, np.zeros(pool_len)))
    for name, Pc, it, s_pool, d_pool in lists:
        ctx = solidfmm_amd.Context(Pc)
        st = solidfmm_amd.Staging(ctx)
        hs, hd = st.reserve(len(s_pool), len(d_pool))
        hs[:] = s_pool
        hd[:] = d_pool
        st.translate(W.M2L, it, len(s_pool), len(d_pool))  # warm: buffers grown
        t = []
        for _ in range(max(args.steps // 50, 3)):
            t0 = time.perf_counter()
            st.translate(W.M2L, it, len(s_pool), len(d_pool))
            t.append(time.perf_counter() - t0)
        pinned_s = float(np.median(t))
        dpool = d_pool.copy()
        ctx.translate_host(W.M2L, it, s_pool, dpool)
        t = []
        for _ in range(max(args.steps // 50, 3)):
            t0 = time.perf_counter()
            ctx.translate_host(W.M2L, it, s_pool, dpool)
            t.append(time.perf_counter() - t0)
        pageable_s = float(np.median(t))
        moved = 8 * (len(s_pool) + 2 * len(d_pool))
        rows.append({"list": name, "items": int(len(it)), "pcie_bytes": moved,
                     "pinned": {"s": pinned_s, "translations_per_s": len(it) / pinned_s,
                                "GB_per_s": moved / pinned_s / 1e9},
                     "pageable": {"s": pageable_s, "translations_per_s": len(it) / pageable_s}})
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
        del st, ctx
    line = {"metric": "host-memory path: translations/s including PCIe copies (sfmm_staging / translate_host)",
            "value": rows[0]["pageable"]["translations_per_s"], "unit": "translations/s", "n_gpus": 1,
            "steps": max(args.steps // 50, 3), "warmup": 1, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "host pools: " + "; ".join(r["list"] for r in rows)}, "results": rows}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--op", choices=list(OPS), default="m2l")
    ap.add_argument("--n", type=int, default=100_000)
    ap.add_argument("--P", type=int, default=10)
    ap.add_argument("--kernels", choices=list(KERNELS), default="auto")
    ap.add_argument("--latency-reps", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--mode", choices=["translate", "mixed", "precision", "serving", "fmm", "host", "copy", "arith"],
                    default="translate",
                    help="mixed: the mixed-order M2L workload (orders --pmin..--pmax, default 500,000 items) "
                         "as one captured HIP graph; copy: time a plain device copy of --copy-mib MiB (HBM "
                         "reference and the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--orders", default="10,16,24", help="orders of the precision study")
    ap.add_argument("--points", type=int, default=1_000_000, help="serving: Plummer points")
    ap.add_argument("--load", type=float, default=0.8,
                    help="serving: Poisson arrival rate as a fraction of the measured one-at-a-time capacity")
    ap.add_argument("--slots", type=int, default=6)
    ap.add_argument("--budget-mib", type=int, default=1024)
    ap.add_argument("--precisions", default="fp64,split32")
    ap.add_argument("--pmin", type=int, default=4)
    ap.add_argument("--pmax", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--theta", type=float, default=0.6, help="fmm/serving: acceptance parameter of the MAC lists")
    ap.add_argument("--server", choices=["native", "python"], default="native",
                    help="serving: the library's native request loop or the Python one")
    ap.add_argument("--lists", choices=["mac", "vlist"], default="mac",
                    help="serving: dual-tree MAC lists or same-level V-lists")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    if args.mode == "copy":
        return copy_mode(args)
    if args.mode == "arith":
        return arith_mode(args)
    if args.mode == "mixed":
        return mixed_mode(args)
    if args.mode == "precision":
        return precision_mode(args)
    if args.mode == "fmm":
        return fmm_mode(args)
    if args.mode == "host":
        return host_mode(args)
    if args.mode == "serving":
        return serving_mode(args)

    import torch

    import solidfmm_amd
    from solidfmm_amd import workload as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    solidfmm_amd.load()
    P, n, op = args.P, args.n, args.op
    seed = 0x5011DF00
    items, src, dst = W.uniform_batch(OPS[op], P, P, n, seed, shift="m2l" if op == "m2l" else "octant")
    ctx = solidfmm_amd.Context(P)
    ctx.set_kernels(KERNELS[args.kernels])
    plan = solidfmm_amd.Plan(ctx, OPS[op], items)
    assert plan.launches == 1
    s = torch.from_numpy(src).cuda()
    d = torch.zeros(n * P * (P + 1), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()

    # One pass on zeroed targets for the error check against the reference.
    plan.execute(s, d, stream)
    torch.cuda.synchronize()
    gpu_out = d.cpu().numpy()

    for _ in range(args.warmup):
        plan.execute(s, d, stream)
    torch.cuda.synchronize()

    # Timed region: K back-to-back steps between two events on the launch
    # stream (per-step events would add their own gaps).
    K = args.steps
    e_start, e_stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_start.record(stream)
    for k in range(K):
        plan.execute(s, d, stream)
    e_stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = e_start.elapsed_time(e_stop)
    # Step intervals with an event between consecutive steps (separate pass).
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(min(K, 200) + 1)]
    ev[0].record(stream)
    for k in range(len(ev) - 1):
        plan.execute(s, d, stream)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(len(ev) - 1)])

    # Isolated batch latency: launch, wait, repeat.
    lat = []
    for _ in range(args.latency_reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        plan.execute(s, d, stream)
        b.record(stream)
        b.synchronize()
        lat.append(a.elapsed_time(b))
    lat = np.array(lat) if lat else np.zeros(1)

    ms_per_step = total_ms / K
    value = n / (ms_per_step * 1e-3)
    bpt = algorithmic_bytes(P, P)
    fpt = algorithmic_flops(op, P)
    achieved = value * bpt / 1e9  # GB/s; one launch per step, so also per launch
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{op}_p{P}_n{n}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    headline = op == "m2l"
    metric = ("M2L translations/s at P=10 FP64 on 1 GPU + p99 batch latency at 1e-10 rel err" if headline
              else f"{op.upper()} translations/s at P={P} FP64 on 1 GPU")
    line = {
        "metric": metric,
        "value": value,
        "unit": "translations/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{op.upper()} FP64, P={P} (degrees 0..{P - 1}), {n} independent translations "
                               f"per batch, " + ("|shift| in [2,4]" if op == "m2l" else "octant shifts")
                               + ", seeded",
                   "P": P, "batch": n, "op": op, "kernels": args.kernels},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "step_interval_p50": float(np.percentile(step_ms, 50)),
                       "step_interval_p99": float(np.percentile(step_ms, 99))},
        "wall_s": wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": traffic,
                     "algorithmic_bytes_per_translation": bpt,
                     "fp64_tflops": value * fpt / 1e12, "fp64_frac": value * fpt / 78.6e12},
    }
    if not args.no_cpu_baseline:
        base, err = cpu_baseline(op, items, src, gpu_out, P, args.cpu_threads)
        line["cpu_baseline"] = base
        per_item = err.max(axis=1)
        line["rel_err_vs_reference"] = {"p50": float(np.percentile(per_item, 50)),
                                        "p99": float(np.percentile(per_item, 99)),
                                        "max": float(per_item.max())}
        line["rel_err_vs_truth"] = truth_errors(op, items, src, gpu_out, P)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
