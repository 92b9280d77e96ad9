#!/usr/bin/env python3
"""Benchmark of the plain per-query GPU form on one MI355X.

Default workload: the 3-way star join + group-by over 120,000,000 fact rows
(lo_suppkey, lo_partkey, lo_orderdate, lo_revenue; supplier 40,000 rows with
region = 1, part 1,400,000 rows with category = 1, date 2,556 rows; group by
(year, brand) into 7 x 1,000 int64 sums).  `--workload q11` runs the
filter+join+sum over 6,001,215 rows instead.

A step is one whole query with every column already resident in HBM: zero the
three hash tables and the result, build the three perfect-hash tables, run the
fused probe + group-by kernel over the fact table.  W untimed steps, then K
timed steps between two synchronizations.  The fused kernel's launches are
bracketed with HIP events on the stream they run on (torch's current stream),
giving the kernel duration used for the HBM roofline.

The answer of the last step is checked bit-exactly against the oracle (the C
restatement of the reference CPU path) on the same inputs, and the oracle's
full query is timed on the host cores as the CPU baseline.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fact rows/s + % HBM roofline; p99 latency, learned+batched vs plain, 1 GPU"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name: str, rows: int | None):
    from mordred_amd import query as Q
    if name == "q11":
        n = rows or 6_001_215
        q = Q.q11(1993)
        desc = {"workload": "filter+join+sum (SSB Q1.1 shape)", "fact_rows": n,
                "predicates": "d_year=1993, discount 1..3, quantity 0..24",
                "aggregate": "int64 SUM(extendedprice*discount)"}
        return q, Q.make_data(q, n, seed=2024), desc
    n = rows or 120_000_000
    q = Q.q21(40_000, 1_400_000)
    desc = {"workload": "3-way star join + group-by (SSB Q2.1 shape)", "fact_rows": n,
            "dims": {"supplier": 40_000, "part": 1_400_000, "date": 2_556},
            "predicates": "s_region=1 of 5, p_category=1 of 25",
            "group_by": "(d_year, p_brand1) 7x1000 dense int64 SUM(lo_revenue)"}
    return q, Q.make_data(q, n, seed=2024, S=40_000, P=1_400_000), desc


def cpu_baseline(q, data, budget_s: float):
    """Oracle (C + OpenMP restatement of the reference CPU path): whole query
    (builds + fact operator) on the same inputs, repeated within budget_s."""
    import oracle
    from oracle import queries as OQ
    cores = min(16, os.cpu_count() or 1)
    oracle.set_threads(cores)
    times = []
    res = None
    t_end = time.perf_counter() + budget_s
    while True:
        t0 = time.perf_counter()
        res = OQ.run_query(q, data, oracle.CONV_GPU)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > t_end or len(times) >= 5:
            break
    med = statistics.median(times)
    return res, {"value": data.n / med, "unit": "fact rows/s", "cores": cores, "kind": "port",
                 "sample": f"whole query on all {data.n} fact rows (3 builds + fused fact operator), "
                           f"median of {len(times)} runs, {med * 1e3:.1f} ms each"}


def copy_probe():
    """A device copy of exactly 1,920,000,000 bytes read + written, repeated 5
    times: the reference kernel for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = 480_000_000
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 4 * n, "calls": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["q21", "q11"], default="q21")
    ap.add_argument("--rows", type=int, default=None, help="override fact rows")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity check")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1.92 GB device-to-device copy (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this path runs on one GPU; multi-GPU runs are not part of this "
                                   "build (see DESIGN.md for how it would shard)"}))
        sys.exit(2)

    import torch
    from mordred_amd import query as Q
    from mordred_amd import abi

    abi.load()
    t0 = time.time()
    q, data, desc = workload(args.workload, args.rows)
    log(f"generated {data.n} rows in {time.time() - t0:.1f}s")
    cache = Q.load_cache(data, q)
    gq = Q.GpuQuery(q, data, cache)
    torch.cuda.synchronize()
    log("columns resident in HBM")

    for _ in range(args.warmup):
        gq.run()
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        gq.reset()
        gq.build()
        ev[i][0].record(stream)
        gq.probe()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    ms_step = wall * 1e3 / args.steps
    k_avg = sum(kernel_ms) / len(kernel_ms)
    log(f"step {ms_step:.3f} ms, fused kernel avg {k_avg:.3f} ms (min {min(kernel_ms):.3f})")

    ncols = len(q.fact_columns())
    alg_bytes = ncols * 4 * data.n
    achieved = alg_bytes / (k_avg * 1e-3) / 1e9
    gpu_res = gq.result()

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath) and (args.rows is None):
        with open(tpath) as f:
            traffic = json.load(f).get(args.workload, {}).get("bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": data.n / (ms_step * 1e-3),
        "unit": "fact rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 columns, int64 sums",
        "data": "synthetic, seeded SplitMix64 columns (mordred_amd/datagen.py)",
        "config": dict(desc, form="plain per-query GPU (builds + fused fact kernel per step)"),
        "kernel_rows_per_s": data.n / (k_avg * 1e-3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE of this kernel)",
                     "kernel": "pipeline_kernel (fused fact probe+group-by)" if q.kind == "groupby"
                     else "pipeline_kernel (fused fact filter+probe+sum)",
                     "algorithmic_bytes": alg_bytes, "kernel_ms_avg": k_avg},
    }
    if not args.no_cpu:
        log("oracle: parity check + CPU baseline")
        ref, base = cpu_baseline(q, data, args.cpu_budget)
        out["cpu_baseline"] = base
        out["parity"] = bool((ref == gpu_res).all())
    else:
        out["cpu_baseline"] = None
        out["parity"] = None
    print(json.dumps(out), flush=True)
    if out["parity"] is False:
        log("PARITY FAILURE: GPU result differs from the oracle")
        sys.exit(1)


if __name__ == "__main__":
    main()
