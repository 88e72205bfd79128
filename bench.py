#!/usr/bin/env python3
"""Benchmark of the NDJSON index + query path on one MI355X.

Workload (default): 4 GiB of seeded tweet-shaped NDJSON, resident in HBM; one
step runs four plain passes, each "index then query" for one of
$.user.lang, $.id, $.place.country_code, $.user.stats.followers.
value = input bytes processed per second over the K timed steps (GB = 1e9 B).

Prints ONE JSON line.  The roofline object uses the index kernel (the dominant
kernel): achieved = input bytes per launch / mean index-kernel duration from HIP
events on the library's stream.  cpu_baseline times the multithreaded CPU
evaluator (test oracle, kind "port") on a bounded prefix of the same input, on
every CPU the process may use.  Streaming copy and read probes give the
achievable HBM bandwidth beside the 8 TB/s peak.

The same line carries the other two halves of the metric (leave them out with
--no-batch / --no-serve; --mode batch / --mode serve print them as lines of
their own):
  batched  256 files of lognormal size in [1 MiB, 512 MiB] (16 GiB), tweets
           and adversarial alternating, 8 member-chain queries per file: one
           batched hipGraph run versus 256 sequential plain runs, checked bit
           for bit against the plain spans and, on a sample of files, the oracle.
  served   64 tweets files (100 GiB) resident under an HBM budget; closed
           burst capacity, then Poisson open-loop arrivals at 3 rates over 12
           member-chain paths (p50/p99 latency from the scheduled arrival,
           GB/s), versus the plain one-query-one-launch path.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PLAIN_PATHS = ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers"]
HBM_PEAK_GBPS = 8000.0
METRIC = "NDJSON GB/s (index+query) as % of HBM peak; p50/p99 latency, batched vs plain"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(data, paths, sample_bytes: int, threads: int):
    """Multithreaded CPU evaluator (oracle) on the first sample_bytes of data:
    (spans per path, baseline record)."""
    import numpy as np

    from gpjson_amd import compile_path
    from oracle import oracle as O

    n = min(sample_bytes, data.size)
    # cut at a newline so the sample is whole lines
    cut = int(np.flatnonzero(data[:n] == 10)[-1]) + 1 if n < data.size else n
    sample = data[:cut]
    cps = [compile_path(p) for p in paths]
    spans = []
    t0 = time.perf_counter()
    for cp in cps:
        nl, s, lb = O.index_mt(sample, cp.max_depth, threads)
        spans.append(O.find_value_mt(sample, nl, s, lb, cp.max_depth, cp.ir, cp.num_results, threads))
    dt = time.perf_counter() - t0
    return spans, {
        "value": round(len(paths) * sample.size / dt / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "host_cpus_available": len(os.sched_getaffinity(0)),
        "host_cpu_count": os.cpu_count(),
        "kind": "port",
        "sample": f"first {sample.size} bytes of the same input, {len(paths)} index+query passes "
                  f"({', '.join(paths)}), {threads} threads, {dt:.2f} s wall = {dt * threads:.0f} thread-seconds",
    }


def load_traffic(kind: str, size: int, seed: int, paths):
    """HBM bytes per index launch from the committed rocprofv3 PMC summary, when it
    was measured on this same workload (else None)."""
    path = os.path.join(ROOT, "profiles", "index_kernel_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    w = d.get("workload", {})
    if (w.get("kind"), w.get("size"), w.get("seed"), w.get("paths")) != (kind, size, seed, list(paths)):
        return None
    return d.get("bytes_per_launch")


TWEET_PATHS_8 = ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers", "$.user.name",
                 "$.user.stats.friends", "$.retweet_count", "$.user.profile.colors.background"]
ADVERSARIAL_PATHS_8 = ["$.id", "$.s1", "$.chain.pad", "$.chain.c1.pad", "$.chain.c1.c2", "$.chain.c1.c2.c3.c4",
                       "$.deep.d.d", "$.s2"]
SERVE_PATHS_12 = ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers", "$.user.name", "$.user.id",
                  "$.user.stats.friends", "$.user.screen_name", "$.retweet_count", "$.lang",
                  "$.place.bounding_box.type", "$.user.profile.colors.background"]


def common_fields(args, value, dt):
    return {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / max(args.steps, 1) * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
    }


def run_plain(args) -> int:
    import numpy as np

    import gpjson_amd
    from gpjson_amd import compile_path, datagen

    paths = [p for p in args.paths.split(",") if p]
    log(f"generating {args.size} bytes of {args.kind} (seed {args.seed})")
    t = time.perf_counter()
    data = datagen.generate(args.kind, args.seed, args.size)
    log(f"generated in {time.perf_counter() - t:.1f} s")

    ctx = gpjson_amd.Context(0)
    t = time.perf_counter()
    buf = ctx.stage(data)
    ctx.synchronize()
    stage_s = time.perf_counter() - t
    log(f"staged into HBM in {stage_s:.2f} s ({data.size / stage_s / 1e9:.1f} GB/s host->device incl. pinned copy)")

    cps = [compile_path(p) for p in paths]
    if args.config:
        ctx.set_config(*[int(x) for x in args.config.split(",")])
    if args.autotune:
        t = time.perf_counter()
        report = ctx.autotune(buf, [cps[0].ir], 2)
        for line in report.strip().splitlines():
            log("autotune: " + line)
        log(f"autotune took {time.perf_counter() - t:.1f} s")
    if args.fusion:
        ctx.set_fusion(True)
    log(f"launch shapes: {ctx.get_config()}; in-tile evaluation {'on' if args.fusion else 'off'}")
    plans = [ctx.plan(buf, [cp.ir]) for cp in cps]
    lines = plans[0].lines
    log(f"plans ready: {lines} lines, levels {[cp.max_depth for cp in cps]}")

    for _ in range(args.warmup):
        for p in plans:
            p.run()
    ctx.synchronize()

    t0 = time.perf_counter()
    for k in range(args.steps):
        for p in plans:
            p.run(timed=(k == args.steps - 1))
    ctx.synchronize()
    t1 = time.perf_counter()
    dt = t1 - t0
    step_bytes = data.size * len(plans)
    value = args.steps * step_bytes / dt / 1e9

    timings = [p.timing() for p in plans]
    fusion = [p.fused() for p in plans]
    log(f"in-tile evaluation per path (fused, lines left to the interpreter): {fusion}")
    idx_ms = [a for a, _ in timings]
    q_ms = [b for _, b in timings]
    mean_idx_ms = sum(idx_ms) / len(idx_ms)
    achieved = data.size / (mean_idx_ms / 1e3) / 1e9
    log(f"index kernel ms per path {['%.3f' % x for x in idx_ms]}; query kernel ms {['%.3f' % x for x in q_ms]}")

    # every plan's spans equal the one-shot path's (its own index + launch)
    plan_spans = [p.spans()[0] for p in plans]
    one_shot_ok = True
    for cp, got in zip(cps, plan_spans):
        idx = ctx.index(buf, cp.max_depth)
        want = ctx.run_queries(idx, [cp.ir], cp.num_results)[0]
        idx.free()
        same = bool(np.array_equal(got, want))
        one_shot_ok &= same
        log(f"{cp.path}: {int(np.sum(got[:, 0, 0] >= 0))} of {lines} lines have a result; "
            f"equal to the one-shot path: {same}")

    traffic = load_traffic(args.kind, args.size, args.seed, paths)
    result = common_fields(args, value, dt)
    result.update({
        "data": f"synthetic seeded {args.kind} NDJSON (seed {args.seed}), resident in HBM",
        "config": {
            "workload": f"plain path, {args.kind} NDJSON {args.size} bytes ({lines} lines); one step = "
                        f"{len(paths)} index+query passes: {', '.join(paths)}",
            "input_bytes": data.size,
            "paths": paths,
            "levels": [cp.max_depth for cp in cps],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "kernel": "index_kernel",
            "algorithmic_bytes_per_launch": data.size,
            "mean_launch_ms": round(mean_idx_ms, 4),
        },
        "frac_of_hbm_peak_end_to_end": round(value / HBM_PEAK_GBPS, 4),
        "index_ms": [round(x, 4) for x in idx_ms],
        "query_ms": [round(x, 4) for x in q_ms],
        "staging_GBps": round(data.size / stage_s / 1e9, 2),
        "launch_shapes": ctx.get_config(),
        "in_tile_evaluation": [{"fused": f, "interpreted_lines": n} for f, n in fusion],
        "bit_exact_vs_one_shot": one_shot_ok,
    })
    for p in plans:
        p.free()
    buf.free()
    copy_ms = ctx.time_copy(1 << 30, 5)
    read_ms = ctx.time_read(1 << 30, 5)
    ctx.close()
    result["hbm_copy_GBps"] = round(2 * (1 << 30) / (copy_ms / 1e3) / 1e9, 1)
    result["hbm_read_GBps"] = round((1 << 30) / (read_ms / 1e3) / 1e9, 1)
    result["roofline"]["frac_of_copy"] = round(achieved / result["hbm_copy_GBps"], 4)
    result["roofline"]["frac_of_read"] = round(achieved / result["hbm_read_GBps"], 4)
    log(f"streaming copy of 1 GiB: {copy_ms:.3f} ms = {result['hbm_copy_GBps']} GB/s (read + write); "
        f"streaming read: {read_ms:.3f} ms = {result['hbm_read_GBps']} GB/s")
    ok = one_shot_ok
    if not args.no_cpu:
        threads = cpu_threads(args)
        log(f"cpu baseline on {min(args.cpu_sample, data.size)} bytes with {threads} threads")
        cpu_spans, result["cpu_baseline"] = cpu_baseline(data, paths, args.cpu_sample, threads)
        # the sample ends at a newline, so all but its last (empty) line are whole lines of the input
        oracle_ok = True
        for got, ref in zip(plan_spans, cpu_spans):
            k = ref.size // (2 * got.shape[1]) - 1
            oracle_ok &= bool(np.array_equal(got[:k].reshape(-1), ref[:k * got.shape[1] * 2]))
        result["bit_exact_vs_oracle_on_sample"] = oracle_ok
        ok &= oracle_ok
        log(f"GPU spans equal the oracle's on the sample's lines: {oracle_ok}")
    else:
        result["cpu_baseline"] = None
    del data, plan_spans
    if not args.no_batch:
        result["batched"] = batch_section(args)
        ok &= result["batched"]["bit_exact_vs_plain"] and result["batched"]["bit_exact_vs_oracle_sample"]
    if not args.no_serve:
        result["served"] = serve_section(args)
        ok &= result["served"]["bit_exact_vs_plain"]
    print(json.dumps(result), flush=True)
    return 0 if ok else 1


def cpu_threads(args) -> int:
    """Threads for the CPU evaluator: every CPU this process may run on, unless
    --cpu-threads says otherwise."""
    return args.cpu_threads or len(os.sched_getaffinity(0))


def batch_sizes(n: int, total: int, seed: int, lo: int = 1 << 20, hi: int = 512 << 20):
    """n lognormal sizes clipped to [lo, hi], scaled so they sum to about total."""
    import numpy as np
    rng = np.random.default_rng(seed)
    raw = rng.lognormal(mean=0.0, sigma=1.5, size=n)
    scale = total / raw.sum()
    for _ in range(50):  # rescale until the clipped sizes sum to the target
        s = np.clip(raw * scale, lo, hi)
        scale *= total / s.sum()
    return [int(x) for x in np.clip(raw * scale, lo, hi)]


def batch_section(args) -> dict:
    """256 files of uneven size (tweets and adversarial mixed), 8 member-chain queries
    per file: one batched run (persistent index launch + one interpreter launch,
    replayed as a hipGraph, launch shapes from the batch autotuner) against 256
    sequential plain runs (one index + one launch of the file's 8 queries per
    file), and against the drop-in's one-query passes (index + query per
    (file, query) pair, 2048 passes).  Every (file, query) is compared with the
    plain path bit for bit, and a sample of files with the multithreaded oracle."""
    import numpy as np

    import gpjson_amd
    from gpjson_amd import compile_path, datagen
    from oracle import oracle as O

    nfiles = args.batch_files
    sizes = batch_sizes(nfiles, args.batch_total, args.seed)
    kinds = ["tweets" if i % 2 == 0 else "adversarial" for i in range(nfiles)]
    log(f"batch: {nfiles} files, {sum(sizes)} bytes (min {min(sizes)}, max {max(sizes)})")
    ctx = gpjson_amd.Context(0)
    bufs, irs, cps_per_file, nbytes = [], [], [], 0
    # the oracle's sample: the smallest file of each kind plus every 32nd file
    sample = set(range(0, nfiles, 32)) | {int(np.argmin(sizes[0::2])) * 2}
    if nfiles > 1:
        sample.add(int(np.argmin(sizes[1::2])) * 2 + 1)
    sample_data = {}
    t = time.perf_counter()
    for i, (kind, n) in enumerate(zip(kinds, sizes)):
        d = datagen.generate(kind, args.seed * 1000 + i, n)
        bufs.append(ctx.stage(d))
        nbytes += d.size
        if i in sample:
            sample_data[i] = d
        del d
        ps = TWEET_PATHS_8 if kind == "tweets" else ADVERSARIAL_PATHS_8
        cps_per_file.append([compile_path(p) for p in ps])
        irs.append([c.ir for c in cps_per_file[-1]])
        if i % 64 == 63:
            log(f"staged {i + 1} files, {nbytes / 1e9:.1f} GB, {time.perf_counter() - t:.1f} s")
    ctx.synchronize()
    if args.config:
        ctx.set_config(*[int(x) for x in args.config.split(",")])
    if args.autotune:
        big = int(np.argmax(sizes))
        t = time.perf_counter()
        report = ctx.autotune(bufs[big], irs[big], 2)
        for line in report.strip().splitlines():
            log("autotune: " + line)
        log(f"autotune on file {big} took {time.perf_counter() - t:.1f} s")
    shapes = {**ctx.get_config(), "query_order": ctx.get_query_order()}
    log(f"launch shapes: {shapes}")

    # plain: one index + one interpreter launch (8 queries) per file, file after file
    plans = [ctx.plan(b, q) for b, q in zip(bufs, irs)]
    for p in plans:
        p.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        for p in plans:
            p.run()
    ctx.synchronize()
    plain_dt = (time.perf_counter() - t0) / args.batch_steps
    plain_gbps = nbytes / plain_dt / 1e9
    log(f"plain: {plain_dt * 1e3:.2f} ms per pass over all files, {plain_gbps:.1f} GB/s")
    # the drop-in's own plain path, one file and one query per index + query
    # pass: 8 passes per file answer what one pass above does (plans are made
    # outside the timed runs, a file's 8 at a time)
    one_dt = 0.0
    for b, q in zip(bufs, irs):
        ps = [ctx.plan(b, [ir]) for ir in q]
        for p in ps:
            p.run()
        ctx.synchronize()
        t0 = time.perf_counter()
        for p in ps:
            p.run()
        ctx.synchronize()
        one_dt += time.perf_counter() - t0
        for p in ps:
            p.free()
    log(f"plain, one query per pass: {one_dt * 1e3:.2f} ms for every (file, query) pair "
        f"({sum(len(q) for q in irs)} passes)")

    batch = ctx.batch(bufs, irs)
    batch_shapes = {**batch.get_config(), "query_order": batch.query_order()}
    tune_s = None
    if not args.no_batch_autotune:
        # the batch's own shapes (tile = chunk of input per work item, block,
        # interpreter shape, lines per claim), timed on the batch itself
        t = time.perf_counter()
        report = batch.autotune(args.batch_tune_iters)
        tune_s = time.perf_counter() - t
        lines = report.strip().splitlines()
        for line in lines[:-1]:
            log("batch autotune: " + line)
        log(f"{lines[-1]} ({tune_s:.1f} s)")
        batch_shapes = {**batch.get_config(), "query_order": batch.query_order()}
    for _ in range(2):
        batch.run(graph=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(args.batch_steps):
        batch.run(graph=True, timed=(k == args.batch_steps - 1))
    ctx.synchronize()
    batch_dt = (time.perf_counter() - t0) / args.batch_steps
    value = nbytes / batch_dt / 1e9
    _, total_ms = batch.timing()
    log(f"batched: {batch_dt * 1e3:.2f} ms per run, {value:.1f} GB/s (events {total_ms:.2f} ms)")

    # the plain runs again with the batch's tuned shapes, for a like-for-like ratio
    plain_same = None
    if batch_shapes != shapes:
        ctx.set_config(*list(batch_shapes.values())[:7])
        ctx.set_query_order(batch_shapes["query_order"])
        tuned_plans = [ctx.plan(b, q) for b, q in zip(bufs, irs)]
        for p in tuned_plans:
            p.run()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            for p in tuned_plans:
                p.run()
        ctx.synchronize()
        plain_same = nbytes / ((time.perf_counter() - t0) / args.batch_steps) / 1e9
        for p in tuned_plans:
            p.free()
        ctx.set_config(*list(shapes.values())[:7])
        ctx.set_query_order(shapes["query_order"])
        log(f"plain with the batch's shapes: {plain_same:.1f} GB/s")

    # bit-exact: every (file, query) of the batch equals the plain plan
    mism = 0
    for f, p in enumerate(plans):
        ps = p.spans()
        for q in range(len(irs[f])):
            if not np.array_equal(batch.spans(f, q), ps[q]):
                mism += 1
    log(f"batched vs plain: {mism} mismatching (file, query) pairs of {sum(len(q) for q in irs)}")
    # and a sample of files against the multithreaded oracle
    threads = cpu_threads(args)
    o_mism, o_bytes = 0, 0
    for f, d in sorted(sample_data.items()):
        cps = cps_per_file[f]
        L = max(c.max_depth for c in cps)
        nl, s, lb = O.index_mt(d, L, threads)
        for q, c in enumerate(cps):
            want = O.find_value_mt(d, nl, s, lb, L, c.ir, c.num_results, threads).reshape(nl.size, c.num_results, 2)
            if not np.array_equal(batch.spans(f, q), want):
                o_mism += 1
        o_bytes += d.size
    log(f"batched vs oracle on {len(sample_data)} files ({o_bytes} bytes): {o_mism} mismatching (file, query) pairs")
    batch.free()
    for p in plans:
        p.free()
    ctx.close()
    return {
        "workload": f"{nfiles} files, lognormal sizes in [1 MiB, 512 MiB], {nbytes} bytes, alternating "
                    "tweets/adversarial; 8 member-chain queries per file; one run = one hipGraph replay "
                    "(persistent index launch over every file's tiles + one interpreter launch)",
        "input_bytes": nbytes,
        "GBps": round(value, 3),
        "ms_per_run": round(batch_dt * 1e3, 3),
        "frac_of_hbm_peak": round(value / HBM_PEAK_GBPS, 4),
        "plain_GBps": round(plain_gbps, 3),
        "plain_ms_per_pass": round(plain_dt * 1e3, 3),
        "batched_over_plain": round(value / plain_gbps, 3),
        "plain_one_query_ms_per_pass": round(one_dt * 1e3, 3),
        "batched_over_plain_one_query": round(one_dt / batch_dt, 3),
        "plain_GBps_batch_shapes": round(plain_same, 3) if plain_same else round(plain_gbps, 3),
        "batched_over_plain_batch_shapes": round(value / (plain_same or plain_gbps), 3),
        "batch_launch_shapes": batch_shapes,
        "batch_autotune_s": round(tune_s, 1) if tune_s is not None else None,
        "bit_exact_vs_plain": mism == 0,
        "bit_exact_vs_oracle_sample": o_mism == 0,
        "oracle_sample": f"{len(sample_data)} files ({o_bytes} bytes), all 8 queries each, oracle on {threads} threads",
        "steps": args.batch_steps,
        "launch_shapes": shapes,
    }


def run_batch(args) -> int:
    r = batch_section(args)
    result = common_fields(args, r["GBps"], r["ms_per_run"] / 1e3 * args.batch_steps)
    result["steps"] = args.batch_steps
    result["warmup"] = 2
    result["ms_per_step"] = r["ms_per_run"]
    result.update({
        "data": f"synthetic seeded tweets + adversarial NDJSON (seed {args.seed}), resident in HBM",
        "config": {"workload": "batched path: " + r["workload"], "input_bytes": r["input_bytes"],
                   "tweet_paths": TWEET_PATHS_8, "adversarial_paths": ADVERSARIAL_PATHS_8},
    })
    result.update({k: v for k, v in r.items() if k not in ("workload", "input_bytes", "GBps")})
    print(json.dumps(result), flush=True)
    return 0 if r["bit_exact_vs_plain"] and r["bit_exact_vs_oracle_sample"] else 1


def serve_section(args) -> dict:
    """Resident files under an HBM budget; Poisson open-loop queries at 3 rates,
    coalesced into one interpreter launch per batch, against the plain
    one-query-one-launch path (index + query + result readback per request).
    Latencies run from each request's scheduled arrival.  Served spans of the
    requests on the files the plain path ran are compared with its spans."""
    import numpy as np

    import gpjson_amd
    from gpjson_amd import compile_path, datagen
    from gpjson_amd.serve import Server, open_loop

    nfiles = args.serve_files
    per_file = args.serve_total // nfiles
    levels = max(compile_path(p).max_depth for p in SERVE_PATHS_12)
    rng = np.random.default_rng(args.seed)

    # plain one-query-one-launch on a few files first (their HBM is released before serving)
    ctx = gpjson_amd.Context(0)
    sample_files = min(4, nfiles)
    plain_t, plain_b = 0.0, 0
    plain_spans = {}
    for i in range(sample_files):
        d = datagen.generate("tweets", args.seed * 1000 + i, per_file)
        buf = ctx.stage(d)
        for p in SERVE_PATHS_12:
            plan = ctx.plan(buf, [compile_path(p).ir])
            plan.run()
            plain_spans[(f"file{i:03d}", p)] = plan.spans()[0]
            t0 = time.perf_counter()
            for _ in range(2):
                plan.run()
                plan.spans()
            plain_t += time.perf_counter() - t0
            plain_b += 2 * d.size
            plan.free()
        buf.free()
        del d
    plain_gbps = plain_b / plain_t / 1e9
    plain_ms = plain_t / (sample_files * len(SERVE_PATHS_12) * 2) * 1e3
    log(f"plain one-query-one-launch: {plain_ms:.2f} ms per request, {plain_gbps:.1f} GB/s")
    ctx.close()

    srv = Server(args.serve_budget, levels)
    t = time.perf_counter()
    for i in range(nfiles):
        d = datagen.generate("tweets", args.seed * 1000 + i, per_file)
        srv.add(f"file{i:03d}", d)
        del d
        if i % 16 == 15:
            log(f"registered {i + 1} files ({(i + 1) * per_file / 1e9:.1f} GB), {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    for i in range(nfiles):
        srv.admit(f"file{i:03d}")
    log(f"admitted {nfiles} files in {time.perf_counter() - t:.1f} s; stats {srv.stats()}")

    def draw(n):
        fs = rng.integers(0, nfiles, size=n)
        ps = rng.integers(0, len(SERVE_PATHS_12), size=n)
        return [(f"file{int(f):03d}", SERVE_PATHS_12[int(p)]) for f, p in zip(fs, ps)]

    # capacity: everything submitted at once (closed burst), after two untimed
    # bursts of the same size (they grow the server's result slots, which each
    # burst splits differently; a burst that grows a slot's pinned buffers took
    # twice as long)
    for _ in range(2):
        for tk in [srv.submit(k, p) for k, p in draw(args.serve_burst)]:
            srv.result(tk).release()
    # three timed bursts; the median is the capacity
    burst_rates = []
    for _ in range(3):
        burst = draw(args.serve_burst)
        t0 = time.perf_counter()
        tickets = [srv.submit(k, p) for k, p in burst]
        for tk in tickets:
            srv.result(tk).release()  # results read in place in pinned memory
        burst_rates.append(len(burst) / (time.perf_counter() - t0))
    capacity = sorted(burst_rates)[1]
    cap_dt = args.serve_burst / capacity
    cap_gbps = args.serve_burst * per_file / cap_dt / 1e9
    log(f"bursts of {args.serve_burst}: {[round(x) for x in burst_rates]} requests/s, median {capacity:.0f} = "
        f"{cap_gbps:.1f} GB/s; stats {srv.stats()}")

    rates = [float(x) for x in args.serve_rates.split(",")] if args.serve_rates else \
        [round(capacity * f, 1) for f in (0.25, 0.5, 0.8)]
    per_rate = []
    for r in rates:
        n = max(100, int(r * args.serve_seconds))
        reqs = draw(n)
        st0 = srv.stats()
        lat, wall, nb, _ = open_loop(srv, reqs, r, seed=args.seed + int(r))
        st1 = srv.stats()
        launches = st1["launches"] - st0["launches"]
        per_rate.append({
            "rate_per_s": r,
            "requests": n,
            "p50_ms": round(float(np.percentile(lat, 50)) / 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) / 1e3, 3),
            "GBps": round(nb / wall / 1e9, 3),
            "mean_batch": round(n / max(launches, 1), 2),
        })
        log(f"rate {r}/s: {per_rate[-1]}")
    # bit-exactness, outside the timed runs: every (file, path) the plain path
    # ran, submitted amid requests on every other file and coalesced with them
    mix = draw(args.serve_burst)
    checks = list(plain_spans)
    for k, (key, p) in enumerate(checks):
        mix.insert((k * len(mix)) // len(checks), (key, p))
    tickets = [srv.submit(key, p) for key, p in mix]
    checked = mism = 0
    for tk, (key, p) in zip(tickets, mix):
        with srv.result(tk) as res:
            if (key, p) in plain_spans:
                checked += 1
                mism += 0 if np.array_equal(res.spans, plain_spans[(key, p)]) else 1
    st = srv.stats()
    srv.close()
    log(f"served vs plain: {mism} mismatching of {checked} requests on the plain path's files, "
        f"in a burst of {len(mix)}")
    return {
        "workload": f"{nfiles} tweets files of {per_file} bytes ({nfiles * per_file} bytes), index levels {levels}, "
                    f"HBM budget {args.serve_budget} bytes; 12 member-chain paths over uniformly chosen files; "
                    "Poisson open-loop arrivals, latency from each request's scheduled arrival to its spans "
                    "being host-visible",
        "input_bytes": nfiles * per_file,
        "hbm_budget_bytes": args.serve_budget,
        "capacity_requests_per_s": round(capacity, 1),
        "capacity_bursts_requests_per_s": [round(x, 1) for x in burst_rates],
        "capacity_GBps": round(cap_gbps, 3),
        "rates": per_rate,
        "plain_GBps": round(plain_gbps, 3),
        "plain_ms_per_request": round(plain_ms, 3),
        "served_over_plain": round(cap_gbps / plain_gbps, 3),
        "bit_exact_vs_plain": mism == 0,
        "bit_exact_checked_requests": checked,
        "server_stats": st,
    }


def run_serve(args) -> int:
    r = serve_section(args)
    value = r["capacity_GBps"]
    result = common_fields(args, value, args.serve_burst / r["capacity_requests_per_s"])
    result["steps"] = 1  # one timed closed burst (after one untimed burst)
    result["warmup"] = 1
    result["ms_per_step"] = round(args.serve_burst / r["capacity_requests_per_s"] * 1e3, 4)
    result.update({
        "data": f"synthetic seeded tweets NDJSON (seed {args.seed}), registered in host memory, resident in HBM",
        "config": {"workload": "serving: " + r["workload"] + "; value = closed-burst capacity in input GB/s",
                   "paths": SERVE_PATHS_12, "hbm_budget_bytes": r["hbm_budget_bytes"]},
    })
    result.update({k: v for k, v in r.items() if k not in ("workload", "hbm_budget_bytes")})
    print(json.dumps(result), flush=True)
    return 0 if r["bit_exact_vs_plain"] else 1


def run_copy(args) -> int:
    """Streaming device copy of --size bytes: the achievable HBM bandwidth, and the
    known-traffic kernel (reads and writes --size bytes per launch) used to
    calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    import gpjson_amd
    ctx = gpjson_amd.Context(0)
    ms = ctx.time_copy(args.size, max(args.steps, 1))
    gbps = 2 * args.size / (ms / 1e3) / 1e9
    result = common_fields(args, gbps, ms / 1e3 * args.steps)
    result.update({
        "data": "device buffer (0x5a bytes)",
        "config": {"workload": f"streaming copy, {args.size} bytes read and written per launch"},
        "ms_per_copy": round(ms, 4),
        "frac_of_hbm_peak": round(gbps / HBM_PEAK_GBPS, 4),
    })
    print(json.dumps(result), flush=True)
    ctx.close()
    return 0


def run_gather(args) -> int:
    """Result materialization: the interpreter launch followed by the device-side
    gather of every result element's bytes into one buffer and one copy back,
    against copying the spans back and slicing each element out of the input on
    the host (what the reference does from its second map of the file)."""
    import gpjson_amd
    from gpjson_amd import compile_path, datagen
    paths = [p for p in args.paths.split(",") if p]
    data = datagen.generate(args.kind, args.seed, args.size)
    host = data.tobytes()
    ctx = gpjson_amd.Context(0)
    buf = ctx.stage(data)
    per = []
    total_elems = total_bytes = 0
    dev_s = host_s = 0.0
    for p in paths:
        cp = compile_path(p)
        idx = ctx.index(buf, cp.max_depth)
        irs = [cp.ir]
        ctx.query_gather(idx, irs, cp.num_results).free()  # warm-up
        t = time.perf_counter()
        g = ctx.query_gather(idx, irs, cp.num_results)
        dev = time.perf_counter() - t
        tm = g.timing()
        vals = g.values(0)
        t = time.perf_counter()
        spans = ctx.run_queries(idx, irs, cp.num_results)[0]
        ref = [[None if a < 0 else host[a:max(a, b)] for a, b in row] for row in spans]
        hst = time.perf_counter() - t
        if vals != ref:
            raise SystemExit(f"gathered values differ from the host slices for {p}")
        elems = int((spans[..., 0] >= 0).sum())
        total_elems += elems
        total_bytes += len(g.bytes)
        dev_s += dev
        host_s += hst
        per.append({"path": p, "elements": elems, "bytes": len(g.bytes), "query_ms": round(tm["query_ms"], 4),
                    "gather_ms": round(tm["gather_ms"], 4), "copy_ms": round(tm["copy_ms"], 4),
                    "device_gather_s": round(dev, 4), "spans_then_host_slices_s": round(hst, 4)})
        log(f"{p}: {elems} elements, {len(g.bytes)} bytes; query {tm['query_ms']:.3f} ms, gather "
            f"{tm['gather_ms']:.3f} ms, copy {tm['copy_ms']:.3f} ms; end to end {dev * 1e3:.1f} ms against "
            f"{hst * 1e3:.1f} ms for spans + host slices")
        g.free()
        idx.free()
    rate = total_elems / dev_s
    result = {
        "metric": "query result elements materialized per second (device gather vs host slices)",
        "value": round(rate, 1),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(dev_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": f"synthetic seeded {args.kind} NDJSON (seed {args.seed}), resident in HBM",
        "config": {"workload": f"result gather, {args.kind} NDJSON {args.size} bytes, one query per pass: "
                               + ", ".join(paths)},
        "host_slices_elements_per_s": round(total_elems / host_s, 1),
        "gather_over_host": round(host_s / dev_s, 3),
        "gathered_bytes": total_bytes,
        "per_path": per,
        "bit_exact_vs_host_slices": True,
    }
    print(json.dumps(result), flush=True)
    buf.free()
    ctx.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="plain", choices=["plain", "batch", "serve", "copy", "gather"])
    ap.add_argument("--size", type=int, default=4 << 30, help="plain: input bytes (default 4 GiB)")
    ap.add_argument("--kind", default="tweets", choices=["tweets", "adversarial"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--paths", default=",".join(PLAIN_PATHS))
    ap.add_argument("--cpu-sample", type=int, default=1 << 30)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true", help="search launch shapes on this input first")
    ap.add_argument("--fusion", action="store_true",
                    help="plain: evaluate the lines inside each index tile in the index kernel")
    ap.add_argument("--config", default="", help="index_block,index_wpt,query_mode,query_block,query_group,wpc,lpg")
    ap.add_argument("--batch-files", type=int, default=256)
    ap.add_argument("--batch-total", type=int, default=16 << 30)
    ap.add_argument("--batch-steps", type=int, default=3, help="batched and 256-file plain runs timed")
    ap.add_argument("--no-batch-autotune", action="store_true",
                    help="run the batch with the context's shapes instead of its own autotuned ones")
    ap.add_argument("--batch-tune-iters", type=int, default=2, help="timed runs per candidate shape")
    ap.add_argument("--no-batch", action="store_true", help="default mode: leave out the batched section")
    ap.add_argument("--serve-files", type=int, default=64)
    ap.add_argument("--serve-total", type=int, default=100 << 30)
    ap.add_argument("--serve-budget", type=int, default=192 << 30)
    ap.add_argument("--serve-burst", type=int, default=512)
    ap.add_argument("--serve-rates", default="", help="requests/s, comma separated (default: 25/50/80%% of capacity)")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--no-serve", action="store_true", help="default mode: leave out the served section")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the plain, batched and served "
              "paths are single-GPU by design.", file=sys.stderr)
        return 2
    if args.mode == "batch":
        return run_batch(args)
    if args.mode == "serve":
        return run_serve(args)
    if args.mode == "copy":
        return run_copy(args)
    if args.mode == "gather":
        return run_gather(args)
    return run_plain(args)


if __name__ == "__main__":
    sys.exit(main())
