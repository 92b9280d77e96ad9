#!/usr/bin/env python3
"""Benchmark of the query-execution path on one MI355X.

Default workload (the metric's configuration): a Poisson stream of requests
drawn from the seeded 64-query star mix (mordred_amd/mix.py, seed 2024) over a
1,200,000,000-row fact table (7 int32 fact columns, 33.6 GB, all resident in
HBM; supplier 40,000, part 1,400,000, customer 3,000,000 and date 2,556 rows),
served two ways on the same arrivals:
  served  the request loop: continuous batching into the persistent
          shared-scan kernel (hipGraph, autotuned tile shape), launch sizes
          ranked by the learned fp32 cost model once it is trusted, admission
          by an HBM budget of 200 GB;
  plain   first come first served, each query alone through its builds and
          its fused fact kernel.
A step is 64 consecutive requests of the stream: W untimed warm-up steps, then
K timed steps (one continuous stream of 64 K requests at --rate q/s, default
2,000 q/s, which saturates both forms).  `value` is the served form's fact
rows answered per second (requests/s x 1.2e9); the line also carries both
forms' q/s and p50/p99 latency, and whether every request's answer was
identical in both forms.  The roofline is that of the shared-scan launches:
algorithmic bytes (distinct fact columns x 4 B x rows scanned) over the HIP
event time of the launches.  The plain per-query form's own kernel measurement
(the 3-way join + group-by over 120,000,000 rows) is nested under
`plain_query_kernel`.

Other workloads: --workload q21 (3-way join + group-by, 120M rows), q11
(filter+join+sum, 6,001,215 rows), q4way (4-way join + group-by over the
300,941,312 rows that the 600M-row split caches, in one fused call),
q4way-split (segment cache + CPU/GPU split, 600M rows), batch64 (64-query shared scan, 600M rows), serve-sweep
(offered rates 10..2,000 q/s), skip, q21-mixed.

Answers are checked bit-exactly against the oracle (the C restatement of the
reference CPU path), which is also timed on the host cores as the CPU
baseline.  Prints one JSON line.
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
    if name == "q4way":
        # the GPU share of the 600M-row split: its cached segments in one fused call
        n = rows or 300_941_312
        q = Q.q4way(40_000, 1_400_000, 3_000_000)
        desc = {"workload": "4-way star join + group-by (SSB Q3/Q4 shape), all segments in HBM", "fact_rows": n,
                "dims": {"supplier": 40_000, "part": 1_400_000, "customer": 3_000_000, "date": 2_556},
                "predicates": "s_region=1 of 5, p_category=1 of 25",
                "group_by": "(c_nation, s_nation, d_year) 25x25x7 dense int64 SUM(lo_revenue - lo_supplycost)"}
        return q, Q.make_data(q, n, seed=2024, S=40_000, P=1_400_000, Cn=3_000_000), desc
    n = rows or 120_000_000
    q = Q.q21(40_000, 1_400_000)
    desc = {"workload": "3-way star join + group-by (SSB Q2.1 shape)", "fact_rows": n,
            "dims": {"supplier": 40_000, "part": 1_400_000, "date": 2_556},
            "predicates": "s_region=1 of 5, p_category=1 of 25",
            "group_by": "(d_year, p_brand1) 7x1000 dense int64 SUM(lo_revenue)"}
    return q, Q.make_data(q, n, seed=2024, S=40_000, P=1_400_000), desc


def oracle_threads() -> int:
    """Host threads for the oracle: every core this process may run on, capped
    by OMP_NUM_THREADS when the environment sets the box's CPU share."""
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def host_info() -> dict:
    """nproc (the affinity mask), the CPU model string and the NUMA node count."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    numa = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node") and d[4:].isdigit()]) \
        if os.path.isdir("/sys/devices/system/node") else None
    return {"nproc": len(os.sched_getaffinity(0)), "cpu_model": model, "numa_nodes": numa,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(q, data, budget_s: float, min_runs: int = 10):
    """Oracle (C + OpenMP restatement of the reference CPU path): whole query
    (builds + fact operator) on the same inputs, one warm-up run, then at
    least min_runs timed runs (more while within budget_s); median."""
    import oracle
    from oracle import queries as OQ
    cores = oracle_threads()
    oracle.set_threads(cores)
    res = OQ.run_query(q, data, oracle.CONV_GPU)
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < min_runs or time.perf_counter() < t_end:
        t0 = time.perf_counter()
        OQ.run_query(q, data, oracle.CONV_GPU)
        times.append(time.perf_counter() - t0)
        if len(times) >= 100:
            break
    med = statistics.median(times)
    return res, {"value": data.n / med, "unit": "fact rows/s", "cores": cores, "kind": "port",
                 "sample": f"whole query on all {data.n} fact rows (builds + fused fact operator), one warm-up, "
                           f"median of {len(times)} runs, {med * 1e3:.1f} ms each "
                           f"(min {min(times) * 1e3:.1f}, max {max(times) * 1e3:.1f})",
                 "host": host_info()}


def pcie_h2d_bandwidth(nbytes: int = 1 << 30) -> float:
    """Bytes per second of one pinned host-to-device copy (torch)."""
    import torch
    h = torch.empty(nbytes // 4, dtype=torch.int32, pin_memory=True)
    d = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda")
    d.copy_(h, non_blocking=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        d.copy_(h, non_blocking=True)
    torch.cuda.synchronize()
    return 3 * nbytes / (time.perf_counter() - t0)


def split_bench(args):
    """4-way star join + group-by over 600M fact rows with ~50% of the
    segments cached in HBM and the rest in pinned host memory (the segment
    cache + CPU/GPU split).  Prints one JSON line."""
    from mordred_amd import engine as E
    from mordred_amd import query as Q
    import torch
    n = args.rows or 600_000_000
    q = Q.q4way(40_000, 1_400_000, 3_000_000)
    t0 = time.time()
    data = Q.make_data(q, n, seed=2024, S=40_000, P=1_400_000, Cn=3_000_000)
    log(f"generated {n} rows x {len(q.fact_columns())} columns in {time.time() - t0:.1f}s")
    nseg = (n + (1 << 20) - 1) >> 20
    cache = int(0.5 * nseg + 1) * len(q.fact_columns()) + 64 + 2 * 4 * len(q.fact_columns())
    threads = oracle_threads()
    eng = E.Engine(device=0, cache_segments=cache, threads=threads)
    eng.add_data(data, q)
    eng.place(q, 0.5)
    log("placed 50% of the fact segments in HBM")
    # calibrate the analytic model's two bandwidths on this machine: PCIe from
    # a pinned copy, the CPU constant so that its per-segment estimate matches
    # the measured CPU time per segment of a residency run
    bw_pci = pcie_h2d_bandwidth()
    res, st = eng.run(q, E.PLACE_RESIDENCY)
    cpu_ms_seg = st["cpu_ms"] / max(1, st["cpu_segments"])
    bw_cpu = 42e9 * st["est_cpu_ms_per_segment"] / cpu_ms_seg
    eng.set_bandwidth(bw_cpu, bw_pci)
    runs = {}
    for name, pl in (("residency", E.PLACE_RESIDENCY), ("cost_split_whole", E.PLACE_COST_SPLIT),
                     ("cost_split", E.PLACE_COST_SPLIT), ("gpu_only", E.PLACE_GPU_ONLY),
                     ("cpu_only", E.PLACE_CPU_ONLY)):
        # cost_split_whole: the split streams whole segments only
        eng.set_prefix_streaming(name != "cost_split_whole")
        times, last = [], None
        for _ in range(args.steps):
            r, st = eng.run(q, pl)
            times.append(st["total_ms"])
            last = (r, st)
        runs[name] = {"ms_median": statistics.median(times), "stats": last[1],
                      "rows_per_s": n / (statistics.median(times) * 1e-3), "res": last[0]}
        log(f"{name}: {statistics.median(times):.1f} ms  {last[1]}")
    # learned placement: train online (exploring random candidate plans), then
    # let the model rank the plans while it beats the analytic model
    eng.set_learned_placement(E.LEARNED_EXPLORE)
    for _ in range(args.learn_runs):
        eng.run(q, E.PLACE_COST_SPLIT)
    eng.set_learned_placement(E.LEARNED_USE)
    times, last, chose = [], None, {0: 0, 1: 0}
    for _ in range(args.steps):
        r, st = eng.run(q, E.PLACE_COST_SPLIT)
        times.append(st["total_ms"])
        chose[st["plan_model"]] = chose.get(st["plan_model"], 0) + 1
        last = (r, st)
    runs["cost_split_learned"] = {"ms_median": statistics.median(times), "stats": last[1],
                                  "rows_per_s": n / (statistics.median(times) * 1e-3), "res": last[0]}
    learned = {"train_runs": args.learn_runs, "model": eng.placement_model_stats(),
               "plans_by_learned_model": chose.get(1, 0), "plans_by_analytic_model": chose.get(0, 0),
               "streamed_segments": last[1]["plan_streamed"], "predicted_ms": last[1]["plan_predicted_ms"],
               "analytic_ms": last[1]["plan_analytic_ms"]}
    log(f"cost_split_learned: {statistics.median(times):.1f} ms  {learned}")
    eng.set_learned_placement(E.LEARNED_OFF)
    learned_order = learned_order_phase(eng, q, nseg, args.steps, args.learn_runs)
    best = min(runs, key=lambda k: runs[k]["ms_median"])
    out = {
        "metric": METRIC + " (segment cache + CPU/GPU split)",
        "value": runs[best]["rows_per_s"],
        "unit": "fact rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": runs[best]["ms_median"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 columns, int64 sums",
        "data": "synthetic, seeded SplitMix64 columns (mordred_amd/datagen.py)",
        "config": {"workload": "4-way star join + group-by with segment caching", "fact_rows": n,
                   "segments": nseg, "cached_fraction": 0.5,
                   "group_by": "(c_nation, s_nation, d_year) 25x25x7 int64 SUM(revenue - supplycost)",
                   "placement": best, "cpu_threads": threads},
        "calibrated_bandwidth": {"cpu_Bps": bw_cpu, "pcie_h2d_Bps": bw_pci},
        "learned_placement": learned,
        "learned_order": learned_order,
        "placements": {k: {"ms": v["ms_median"], "rows_per_s": v["rows_per_s"],
                           "stats": {a: b for a, b in v["stats"].items()}} for k, v in runs.items()},
    }
    if not args.no_cpu:
        import oracle
        from oracle import queries as OQ
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        ref = OQ.run_query(q, data, oracle.CONV_GPU)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / cpu_s, "unit": "fact rows/s", "cores": threads, "kind": "port",
                               "sample": f"whole query on all {n} fact rows, one run, {cpu_s * 1e3:.0f} ms"}
        ok = {k: bool((ref == v["res"]).all()) for k, v in runs.items()}
        if not ok.get("cpu_only", True):
            # CPU operators alone follow the CPU path's convention
            ok["cpu_only"] = bool((OQ.run_query(q, data, oracle.CONV_CPU) == runs["cpu_only"]["res"]).all())
        out["parity"] = all(ok.values())
        out["parity_by_placement"] = ok
    print(json.dumps(out), flush=True)
    eng.close()
    if out.get("parity") is False:
        log("PARITY FAILURE")
        sys.exit(1)


def batch_bench(args):
    """64 concurrent star queries of the seeded mix over the 600M-row fact
    table, all columns resident in HBM: one shared-scan batch (hipGraph, tile
    shape from the autotuner) against the plain per-query GPU form (each query
    alone: its builds and its fused kernel).  Prints one JSON line."""
    from mordred_amd import engine as E
    from mordred_amd import mix
    n = args.rows or 600_000_000
    dims = {"S": 40_000, "P": 1_400_000, "C": 3_000_000}
    mq = mix.make_mix(64, 2024, dims)
    qs = [m.query for m in mq]
    t0 = time.time()
    data = mix.make_mix_data(qs, n, seed=2024, S=dims["S"], P=dims["P"], Cn=dims["C"])
    cols = mix.fact_columns(qs)
    log(f"generated {n} rows x {len(cols)} fact columns in {time.time() - t0:.1f}s")
    nseg = (n + (1 << 20) - 1) >> 20
    threads = oracle_threads()
    eng = E.Engine(device=0, cache_segments=len(cols) * nseg + 64, threads=threads)
    t0 = time.time()
    eng.add_mix_data(data, qs)
    log(f"loaded into the segment cache in {time.time() - t0:.1f}s")
    if args.batch_config:
        best = tuple(int(x) for x in args.batch_config.split(","))
        table = []
        log(f"launch configuration {best} (autotuner skipped)")
    else:
        best, table = eng.batch_tune(qs, reps=2)
        log(f"autotuner: best {best}; " + ", ".join(f"{c}:{ms:.2f}" for c, ms in table))
    for _ in range(args.warmup):
        eng.batch(qs, best)
    tot, gpu = [], []
    outs = st = None
    for _ in range(args.steps):
        outs, st = eng.batch(qs, best)
        tot.append(st["total_ms"])
        gpu.append(st["gpu_ms"])
    batch_ms = statistics.median(tot)
    gpu_ms = statistics.median(gpu)
    log(f"batch: {batch_ms:.2f} ms per 64 queries (graph {gpu_ms:.2f} ms), {st}")
    # plain form: every query alone through the per-query path
    plain = []
    same = True
    for rep in range(2):
        t0 = time.perf_counter()
        for q, r in zip(qs, outs):
            res, _ = eng.run(q, E.PLACE_RESIDENCY)
            same = same and bool((res == r).all())
        plain.append((time.perf_counter() - t0) * 1e3)
    plain_ms = min(plain)
    log(f"plain: {plain_ms:.1f} ms per 64 queries; batch == plain: {same}")
    nbytes = 4 * len(cols) * n
    out = {
        "metric": METRIC + " (64-query shared-scan batch)",
        "value": 64 * n / (batch_ms * 1e-3),
        "unit": "fact rows/s summed over the 64 queries",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batch_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 columns, int64 sums",
        "data": "synthetic, seeded SplitMix64 columns (mordred_amd/datagen.py); query mix mordred_amd/mix.py seed 2024",
        "config": {"workload": "64 concurrent star queries, one shared fact scan", "fact_rows": n,
                   "fact_columns": len(cols), "queries": 64,
                   "joins": [len(q.joins) for q in qs], "groups": [q.total_val for q in qs],
                   "expected_selectivity_sum": sum(m.expected_selectivity for m in mq),
                   "launch": {"block": best[0], "items_per_thread": 4 * best[1], "lds_bytes": best[2]}},
        "queries_per_s": 64 / (batch_ms * 1e-3),
        "plain_queries_per_s": 64 / (plain_ms * 1e-3),
        "speedup_vs_plain": plain_ms / batch_ms,
        "roofline": batch_roofline(nbytes, gpu_ms, n, best),
        "batch_stats": st,
        "autotune": [{"config": c, "gpu_ms": ms} for c, ms in table],
        "parity_batch_vs_plain": same,
    }
    if not args.no_cpu:
        import oracle
        from oracle import queries as OQ
        oracle.set_threads(threads)
        # every query of the batch against the oracle over all n rows
        bad = []
        times = []
        for i in range(len(qs)):
            t0 = time.perf_counter()
            ref = OQ.run_query(qs[i], data, oracle.CONV_GPU)
            times.append(time.perf_counter() - t0)
            if not bool((ref == outs[i]).all()):
                bad.append(i)
        cpu_s = statistics.median(times)
        out["cpu_baseline"] = {"value": n / cpu_s, "unit": "fact rows/s per query", "cores": threads, "kind": "port",
                               "sample": f"oracle on each of the {len(qs)} queries over all {n} rows, median "
                                         f"{cpu_s * 1e3:.0f} ms per query (min {min(times) * 1e3:.0f}, max "
                                         f"{max(times) * 1e3:.0f})",
                               "host": host_info()}
        out["parity_oracle_queries"] = len(qs) - len(bad)
        out["parity_oracle_mismatches"] = bad
        ok = not bad
        out["parity"] = ok and same
    else:
        out["parity"] = same
    print(json.dumps(out), flush=True)
    eng.close()
    if not out["parity"]:
        log("PARITY FAILURE")
        sys.exit(1)


def serve_bench(args):
    """Poisson request streams of the 64-query mix over the 1.2B-row fact table
    (all columns in HBM, admission budget 200 GB), rate swept 10..2,000 q/s:
    the plain form (first come first served, each query alone through its
    builds and fused kernel) against continuous batching (the circular shared
    scan, launch sizes ranked by the learned cost model once it is trusted).
    Every request's answer must be identical in both forms.  One JSON line."""
    from mordred_amd import engine as E
    from mordred_amd import mix, serving
    n = args.rows or 1_200_000_000
    dims = {"S": 40_000, "P": 1_400_000, "C": 3_000_000}
    mq = mix.make_mix(64, 2024, dims)
    qs = [m.query for m in mq]
    t0 = time.time()
    data = mix.make_mix_data(qs, n, seed=2024, S=dims["S"], P=dims["P"], Cn=dims["C"])
    cols = mix.fact_columns(qs)
    log(f"generated {n} rows x {len(cols)} fact columns in {time.time() - t0:.1f}s")
    nseg = (n + (1 << 20) - 1) >> 20
    threads = oracle_threads()
    eng = E.Engine(device=0, cache_segments=len(cols) * nseg + 64, threads=threads)
    t0 = time.time()
    eng.add_mix_data(data, qs)
    log(f"loaded into the segment cache in {time.time() - t0:.1f}s")
    best, _ = eng.batch_tune(qs, reps=1)
    log(f"batch configuration {best}")
    budget = 200e9
    target = args.target_chunk_ms
    # warm-up stream: compiles the graph and trains the cost model online
    warm = serving.request_stream(qs, 400, seed=1)
    _, _, wst = eng.serve(warm, serving.poisson_arrivals(4000.0, len(warm), seed=1), E.SERVE_CONTINUOUS,
                          target_chunk_ms=target, hbm_budget_bytes=budget, learned=True)
    log(f"warm-up: {wst}")
    rates = [float(r) for r in args.rates.split(",")]
    rows, same_all = [], True
    for i, rate in enumerate(rates):
        nreq = int(min(600, max(60, rate * 2.0)))
        reqs = serving.request_stream(qs, nreq, seed=100 + i)
        arr = serving.poisson_arrivals(rate, nreq, seed=100 + i)
        po, pt, pst = eng.serve(reqs, arr, E.SERVE_PLAIN)
        co, ct, cst = eng.serve(reqs, arr, E.SERVE_CONTINUOUS, target_chunk_ms=target, hbm_budget_bytes=budget,
                                learned=True, lone_plain=True)
        same = all(bool((a == b).all()) for a, b in zip(po, co))
        same_all = same_all and same
        p, c = serving.latency_summary(pt), serving.latency_summary(ct)
        rows.append({"rate_qps": rate, "requests": nreq, "plain": p, "served": c, "identical": same,
                     "served_launches": cst["launches"], "served_max_active": cst["max_active_seen"],
                     "learned_choices": cst["learned_choices"], "analytic_choices": cst["analytic_choices"],
                     "learned_error": cst["learned_error"], "analytic_error": cst["analytic_error"],
                     "learned_trusted": cst["learned_trusted"], "served_lone_requests": cst["lone_requests"]})
        log(f"rate {rate:7.0f} q/s: plain {p['throughput_qps']:7.1f} q/s p50 {p['p50_ms']:8.1f} p99 {p['p99_ms']:8.1f} | "
            f"served {c['throughput_qps']:7.1f} q/s p50 {c['p50_ms']:8.1f} p99 {c['p99_ms']:8.1f} | "
            f"launches {cst['launches']} lone {cst['lone_requests']} learned {cst['learned_choices']}/{cst['launches']} "
            f"identical {same}")
    top = rows[-1]
    out = {
        "metric": METRIC + " (request loop, Poisson arrivals)",
        "value": top["served"]["throughput_qps"],
        "unit": "queries/s at the highest offered rate",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 1,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 columns, int64 sums",
        "data": "synthetic, seeded SplitMix64 columns (mordred_amd/datagen.py); query mix mordred_amd/mix.py seed 2024",
        "config": {"workload": "Poisson stream of the 64-query star mix, continuous batching vs plain per-query",
                   "fact_rows": n, "fact_columns": len(cols), "hbm_budget_bytes": budget,
                   "target_chunk_ms": target,
                   "launch": {"block": best[0], "items_per_thread": 4 * best[1], "lds_bytes": best[2]}},
        "throughput_vs_plain": top["served"]["throughput_qps"] / top["plain"]["throughput_qps"],
        "p99_vs_plain": top["served"]["p99_ms"] / top["plain"]["p99_ms"],
        "sweep": rows,
        "parity_plain_vs_served": same_all,
    }
    if not args.no_cpu:
        import oracle
        from oracle import queries as OQ
        oracle.set_threads(threads)
        sample = [0, 21, 42]
        ok = True
        t0 = time.perf_counter()
        for j in sample:
            ref = OQ.run_query(qs[j], data, oracle.CONV_GPU)
            got, _, _ = eng.serve([qs[j]], [0.0], E.SERVE_CONTINUOUS, target_chunk_ms=target)
            ok = ok and bool((ref == got[0]).all())
        cpu_s = (time.perf_counter() - t0) / len(sample)
        out["cpu_baseline"] = {"value": 1.0 / cpu_s, "unit": "queries/s", "cores": threads, "kind": "port",
                               "sample": f"oracle on mix queries {sample} over all {n} rows, {cpu_s * 1e3:.0f} ms per query"}
        out["parity_oracle_sample"] = ok
        out["parity"] = ok and same_all
    else:
        out["parity"] = same_all
    print(json.dumps(out), flush=True)
    eng.close()
    if not out["parity"]:
        log("PARITY FAILURE")
        sys.exit(1)


def serve_headline(args):
    """The metric's configuration: one Poisson request stream of the 64-query
    mix over the 1.2B-row fact table, served (continuous batching into the
    shared-scan kernel, learned cost model, HBM admission) and plain (each
    request alone), on the same arrivals.  A step is 64 consecutive requests;
    W warm-up steps, then K timed steps.  Prints one JSON line."""
    import gc

    import torch
    from mordred_amd import engine as E
    from mordred_amd import mix, serving
    n = args.rows or 1_200_000_000
    dims = {"S": 40_000, "P": 1_400_000, "C": 3_000_000}
    mq = mix.make_mix(64, 2024, dims)
    qs = [m.query for m in mq]
    t0 = time.time()
    data = mix.make_mix_data(qs, n, seed=2024, S=dims["S"], P=dims["P"], Cn=dims["C"])
    cols = mix.fact_columns(qs)
    log(f"generated {n} rows x {len(cols)} fact columns in {time.time() - t0:.1f}s")
    nseg = (n + (1 << 20) - 1) >> 20
    threads = oracle_threads()
    eng = E.Engine(device=0, cache_segments=len(cols) * nseg + 64, threads=threads)
    t0 = time.time()
    eng.add_mix_data(data, qs)
    log(f"loaded into the segment cache in {time.time() - t0:.1f}s")
    best, table = eng.batch_tune(qs, reps=1)
    log(f"autotuner: batch configuration {best}")
    budget, target, rate = 200e9, args.target_chunk_ms, args.rate
    # setup: a 400-request stream trains the learned cost model online
    warm = serving.request_stream(qs, 400, seed=1)
    _, _, wst = eng.serve(warm, serving.poisson_arrivals(4000.0, len(warm), seed=1), E.SERVE_CONTINUOUS,
                          target_chunk_ms=target, hbm_budget_bytes=budget, learned=True)
    log(f"cost-model training stream: {wst}")
    nw, nk = 64 * args.warmup, 64 * args.steps
    reqs = serving.request_stream(qs, nw + nk, seed=7)
    arr = serving.poisson_arrivals(rate, nw + nk, seed=7)
    timed, tarr = reqs[nw:], arr[nw:] - arr[nw]

    def run(mode):
        kw = {} if mode == E.SERVE_PLAIN else dict(target_chunk_ms=target, hbm_budget_bytes=budget, learned=True,
                                                   lone_plain=True)
        if nw:
            eng.serve(reqs[:nw], arr[:nw], mode, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        outs, times, st = eng.serve(timed, tarr, mode, **kw)
        torch.cuda.synchronize()
        return outs, times, st, time.perf_counter() - t

    co, ct, cst, c_wall = run(E.SERVE_CONTINUOUS)
    c = serving.latency_summary(ct)
    log(f"served: {c['throughput_qps']:.1f} q/s p50 {c['p50_ms']:.1f} p99 {c['p99_ms']:.1f} ms, {cst}")
    # the same stream with identical concurrent queries scanned separately
    # (the stream repeats queries of the mix; sharing their sums is part of the
    # served form, reported here without it too)
    eng.set_batch_sharing(False)
    uo, ut, ust, _ = run(E.SERVE_CONTINUOUS)
    eng.set_batch_sharing(True)
    u = serving.latency_summary(ut)
    log(f"served, no sharing: {u['throughput_qps']:.1f} q/s p50 {u['p50_ms']:.1f} p99 {u['p99_ms']:.1f} ms")
    po, pt, pst, p_wall = run(E.SERVE_PLAIN)
    p = serving.latency_summary(pt)
    log(f"plain: {p['throughput_qps']:.1f} q/s p50 {p['p50_ms']:.1f} p99 {p['p99_ms']:.1f} ms")
    identical = all(bool((a == b).all()) for a, b in zip(po, co))
    identical_unshared = all(bool((a == b).all()) for a, b in zip(po, uo))
    scan_ms = cst["gpu_ms"] / max(1, cst["launches"])
    achieved = cst["alg_bytes"] / (cst["gpu_ms"] * 1e-3) / 1e9 if cst["gpu_ms"] > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": c["throughput_qps"] * n,
        "unit": "fact rows/s answered (served requests/s x 1.2e9 fact rows per request)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": c_wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 columns, int64 sums",
        "data": "synthetic, seeded SplitMix64 columns (mordred_amd/datagen.py); query mix mordred_amd/mix.py seed "
                "2024; requests drawn uniformly from the mix, Poisson arrivals (seed 7)",
        "config": {"workload": "Poisson request stream of the 64-query star mix over 1.2B fact rows: learned + "
                               "batched + served vs plain per-query GPU form",
                   "fact_rows": n, "fact_columns": len(cols), "offered_qps": rate,
                   "requests_timed": nk, "requests_warmup": nw, "step": "64 consecutive requests",
                   "dims": {"supplier": dims["S"], "part": dims["P"], "customer": dims["C"], "date": 2556},
                   "hbm_budget_bytes": budget, "target_launch_ms": target,
                   "launch": {"block": best[0], "items_per_thread": 4 * best[1], "lds_bytes": best[2]}},
        "served": dict(c, launches=cst["launches"], gpu_ms=cst["gpu_ms"], wall_ms=cst["wall_ms"],
                       max_active=cst["max_active_seen"], lone_requests=cst["lone_requests"],
                       learned_choices=cst["learned_choices"],
                       analytic_choices=cst["analytic_choices"], learned_trusted=cst["learned_trusted"],
                       learned_error=cst["learned_error"], analytic_error=cst["analytic_error"],
                       shared_query_launches=cst["shared_queries"]),
        "served_without_sharing": dict(u, launches=ust["launches"], gpu_ms=ust["gpu_ms"],
                                       identical_to_plain=identical_unshared,
                                       throughput_vs_plain=u["throughput_qps"] / p["throughput_qps"],
                                       p99_vs_plain=u["p99_ms"] / p["p99_ms"],
                                       note="the same stream, identical concurrent queries each scanned "
                                            "(mordred_engine_set_batch_sharing off)"),
        "plain": dict(p, gpu_ms=pst["gpu_ms"], wall_ms=pst["wall_ms"]),
        "throughput_vs_plain": c["throughput_qps"] / p["throughput_qps"],
        "p50_vs_plain": c["p50_ms"] / p["p50_ms"],
        "p99_vs_plain": c["p99_ms"] / p["p99_ms"],
        "identical_served_vs_plain": identical,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "frac_kind": "effective (algorithmic bytes)",
                     "traffic": None, "profile": "profiles/batch_kernel_serve_1200M.md",
                     "kernel": "batch_kernel (shared scan of the served launches)",
                     "algorithmic_bytes": cst["alg_bytes"] / max(1, cst["launches"]),
                     "kernel_ms_avg": scan_ms, "launches": cst["launches"],
                     "note": "per launch: distinct fact columns of the active queries x 4 B x rows of the "
                             "launch; time: HIP events around each launch's graph (clear + packed-table "
                             "builds when the active set changed + the scan), on the engine's stream"},
    }
    # the oracle on a sample of the timed requests: 10 distinct queries in
    # arrival order, one warm-up run, each over all n rows
    if not args.no_cpu:
        import oracle
        from oracle import queries as OQ
        oracle.set_threads(threads)
        seen, sample = set(), []
        for i, q in enumerate(timed):
            if q.name not in seen:
                seen.add(q.name)
                sample.append(i)
            if len(sample) == 10:
                break
        OQ.run_query(timed[sample[0]], data, oracle.CONV_GPU)
        ok, secs = True, []
        for i in sample:
            t = time.perf_counter()
            ref = OQ.run_query(timed[i], data, oracle.CONV_GPU)
            secs.append(time.perf_counter() - t)
            ok = ok and bool((ref == co[i]).all())
        med = statistics.median(secs)
        out["cpu_baseline"] = {"value": n / med, "unit": "fact rows/s (one query at a time)", "cores": threads,
                               "kind": "port", "queries_per_s": 1.0 / med,
                               "sample": f"oracle on 10 distinct queries of the timed stream ({[timed[i].name for i in sample]}) "
                                         f"over all {n} rows, one warm-up run, median {med * 1e3:.0f} ms per query "
                                         f"(min {min(secs) * 1e3:.0f}, max {max(secs) * 1e3:.0f})",
                               "host": host_info()}
        out["parity_oracle_sample"] = ok
    eng.close()
    del eng, data
    gc.collect()
    # the plain per-query form's own kernel: 3-way join + group-by, 120M rows
    q2, d2, desc2 = workload("q21", None)
    out["plain_query_kernel"] = plain_query_measure(q2, d2, desc2, 20, 5, args.no_cpu, 5.0, "q21")
    out["parity"] = identical and out.get("parity_oracle_sample", True) and \
        out["plain_query_kernel"]["parity"] is not False
    print(json.dumps(out), flush=True)
    if not out["parity"]:
        log("PARITY FAILURE")
        sys.exit(1)


def skip_bench(args):
    """SSB Q1.1, Q1.2 and Q1.3 over a 600M-row lineorder sorted by
    lo_orderdate, read from on-disk column files and minmax files, every fact
    segment cached in HBM; each query with segment skipping off and on.
    Prints one JSON line (value: Q1.1 fact rows/s with skipping)."""
    import shutil
    import tempfile

    import numpy as np
    from mordred_amd import columnar
    from mordred_amd import engine as E
    from mordred_amd import query as Q
    n = args.rows or 600_000_000
    qs = [Q.q11(), Q.q12(), Q.q13()]
    cols = qs[0].fact_columns()
    t0 = time.time()
    data = Q.make_data(qs[0], n, seed=2024, sort_orderdate=True)
    log(f"generated {n} rows x {len(cols)} columns (sorted by lo_orderdate) in {time.time() - t0:.1f}s")
    tmp = tempfile.mkdtemp(prefix="mordred_columns_")
    need = 4 * n * len(cols)
    if shutil.disk_usage(tmp).free < need + (1 << 30):
        shutil.rmtree(tmp, ignore_errors=True)
        raise SystemExit(f"skip workload: {need / 1e9:.1f} GB of column files do not fit in {tmp}")
    nseg = (n + (1 << 20) - 1) >> 20
    threads = oracle_threads()
    eng = E.Engine(device=0, cache_segments=nseg * len(cols) + 64, threads=threads)
    try:
        t0 = time.time()
        columnar.write_table(tmp, {c: data.fact[c] for c in cols}, sorted_table=True)
        log(f"wrote {len(cols)} column files + minmax files in {time.time() - t0:.1f}s")
        t0 = time.perf_counter()
        columnar.load_table(eng, tmp, cols, sorted_table=True)
        load_s = time.perf_counter() - t0
        log(f"loaded the column files and minmax files in {load_s:.2f}s "
            f"({4 * n * len(cols) / load_s / 1e9:.1f} GB/s, page cache)")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    for c in ("d_datekey", "d_year", "d_yearmonthnum"):
        eng.add_column(c, data.tables["date"][c], gpu_resident=True)
    for c in cols:
        for sgm in range(nseg):
            eng.cache_segment(c, sgm)
    runs = {}
    for q in qs:
        for skipping in (False, True):
            eng.set_skipping(skipping)
            for _ in range(args.warmup):
                eng.run(q, E.PLACE_RESIDENCY)
            times, gpu, last = [], [], None
            for _ in range(args.steps):
                r, st = eng.run(q, E.PLACE_RESIDENCY)
                times.append(st["total_ms"])
                gpu.append(st["gpu_ms"])
                last = (r, st)
            key = f"{q.name}_{'skip' if skipping else 'scan'}"
            runs[key] = {"ms": statistics.median(times), "gpu_ms": statistics.median(gpu), "res": last[0],
                         "skipped_segments": last[1]["skipped_segments"], "gpu_rows": last[1]["gpu_rows"]}
            log(f"{key}: {runs[key]['ms']:.3f} ms (GPU part {runs[key]['gpu_ms']:.3f} ms), "
                f"skipped {runs[key]['skipped_segments']} of {nseg} segments")
    head = runs["q11_skip"]
    scanned_bytes = head["gpu_rows"] * 4 * len(cols)
    out = {
        "metric": METRIC + " (sorted lineorder, per-segment min/max skipping)",
        "value": n / (head["ms"] * 1e-3),
        "unit": "fact rows/s (table rows over the Q1.1 query time, skipping on)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 columns, int64 sums",
        "data": "synthetic, seeded SplitMix64 columns (mordred_amd/datagen.py), lineorder sorted by lo_orderdate, "
                "written to and read from column files",
        "config": {"workload": "SSB Q1.1-Q1.3 on a lineorder sorted by lo_orderdate, all segments in HBM",
                   "fact_rows": n, "segments": nseg, "fact_columns": cols},
        "load_GBps": 4 * n * len(cols) / load_s / 1e9,
        "queries": {k: {a: b for a, b in v.items() if a != "res"} for k, v in runs.items()},
        "speedup_by_skipping": {q.name: runs[f"{q.name}_scan"]["ms"] / runs[f"{q.name}_skip"]["ms"] for q in qs},
        "roofline": {"bound": "hbm", "achieved": scanned_bytes / (head["gpu_ms"] * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": scanned_bytes / (head["gpu_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "frac_kind": "effective (algorithmic bytes)", "traffic": None,
                     "note": "bytes of the segments Q1.1 reads (4 columns x 4 B x rows) over the host-timed GPU "
                             "part (date build + fused kernel + result copy)"},
    }
    parity = all(bool((runs[f"{q.name}_scan"]["res"] == runs[f"{q.name}_skip"]["res"]).all()) for q in qs)
    if not args.no_cpu:
        import oracle
        from oracle import queries as OQ
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        refs = {q.name: OQ.run_query(q, data, oracle.CONV_GPU) for q in qs}
        cpu_s = (time.perf_counter() - t0) / len(qs)
        parity = parity and all(bool((refs[q.name] == runs[f"{q.name}_skip"]["res"]).all()) for q in qs)
        out["cpu_baseline"] = {"value": n / cpu_s, "unit": "fact rows/s", "cores": threads, "kind": "port",
                               "sample": f"oracle, Q1.1-Q1.3 each on all {n} rows without skipping, "
                                         f"{cpu_s * 1e3:.0f} ms per query"}
    out["parity"] = parity
    print(json.dumps(out), flush=True)
    eng.close()
    if not parity:
        log("PARITY FAILURE")
        sys.exit(1)


def mixed_bench(args):
    """SSB Q2.1 shape over 600M fact rows with only the supplier and part keys
    cached in HBM: the mixed pipeline (GPU probes into offset lists, CPU
    operators over the survivors) against streaming the missing columns and
    against the CPU operators alone.  Prints one JSON line."""
    import torch
    from mordred_amd import engine as E
    from mordred_amd import query as Q
    n = args.rows or 600_000_000
    q = Q.q21(40_000, 1_400_000)
    t0 = time.time()
    data = Q.make_data(q, n, seed=2024, S=40_000, P=1_400_000)
    log(f"generated {n} rows x {len(q.fact_columns())} columns in {time.time() - t0:.1f}s")
    nseg = (n + (1 << 20) - 1) >> 20
    threads = oracle_threads()
    bw_pci = pcie_h2d_bandwidth()   # torch initialises the device before the engine
    torch.cuda.synchronize()
    eng = E.Engine(device=0, cache_segments=3 * nseg + 64 + 2 * 4 * len(q.fact_columns()), threads=threads)
    eng.add_data(data, q)
    cached = ["lo_suppkey", "lo_partkey"]
    for c in cached:
        for sgm in range(nseg):
            eng.cache_segment(c, sgm)
    log(f"cached {cached} of all {nseg} segments")
    eng.set_bandwidth(42e9, bw_pci)
    runs = {}
    for name, pl in (("mixed_residency", E.PLACE_RESIDENCY), ("cost_split", E.PLACE_COST_SPLIT),
                     ("gpu_only_stream", E.PLACE_GPU_ONLY), ("cpu_only", E.PLACE_CPU_ONLY)):
        times, last = [], None
        for _ in range(args.warmup):
            eng.run(q, pl)
        for _ in range(args.steps if pl != E.PLACE_CPU_ONLY else max(1, args.steps // 3)):
            r, st = eng.run(q, pl)
            times.append(st["total_ms"])
            last = (r, st)
        runs[name] = {"ms_median": statistics.median(times), "stats": last[1], "res": last[0]}
        log(f"{name}: {statistics.median(times):.1f} ms  mixed {last[1]['mixed_segments']} "
            f"survivors {last[1]['mixed_survivors']} stream {last[1]['stream_segments']} cpu {last[1]['cpu_segments']}")
    # operator orders (the reference's executeTableFact_v1 / _v2 beside the
    # GPU-first order), with the keys cached and then also lo_revenue, which
    # lets the CPU-first and alternate orders run the group-by on the GPU
    orders = {}
    for extra in ([], ["lo_revenue"]):
        for c in extra:
            for sgm in range(nseg):
                eng.cache_segment(c, sgm)
        for oname, order in (("gpu_first", E.ORDER_GPU_FIRST), ("cpu_first_v1", E.ORDER_CPU_FIRST),
                             ("alternate_v2", E.ORDER_ALTERNATE)):
            eng.set_operator_order(order)
            for _ in range(args.warmup):
                eng.run(q, E.PLACE_RESIDENCY)
            times, last = [], None
            for _ in range(args.steps):
                r, st = eng.run(q, E.PLACE_RESIDENCY)
                times.append(st["total_ms"])
                last = (r, st)
            key = f"{oname}{'+revenue' if extra else ''}"
            orders[key] = {"ms_median": statistics.median(times), "stats": last[1], "res": last[0]}
            log(f"order {key}: {statistics.median(times):.1f} ms  switches {last[1]['order_switches']} "
                f"survivors {last[1]['mixed_survivors']}")
        eng.set_operator_order(E.ORDER_GPU_FIRST)
    best = min(runs, key=lambda k: runs[k]["ms_median"])
    out = {
        "metric": METRIC + " (partly cached fact table, mixed CPU/GPU operator pipelines)",
        "value": n / (runs["mixed_residency"]["ms_median"] * 1e-3),
        "unit": "fact rows/s (mixed pipeline)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": runs["mixed_residency"]["ms_median"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32 columns, int64 sums",
        "data": "synthetic, seeded SplitMix64 columns (mordred_amd/datagen.py)",
        "config": {"workload": "3-way star join + group-by, supplier and part keys cached, date key and revenue "
                               "in host memory", "fact_rows": n, "segments": nseg, "cached_columns": cached,
                   "cpu_threads": threads, "best": best},
        "placements": {k: {"ms": v["ms_median"], "stats": v["stats"]} for k, v in runs.items()},
        "speedup_mixed_vs_stream": runs["gpu_only_stream"]["ms_median"] / runs["mixed_residency"]["ms_median"],
        "speedup_mixed_vs_cpu": runs["cpu_only"]["ms_median"] / runs["mixed_residency"]["ms_median"],
        "operator_orders": {k: {"ms": v["ms_median"], "switches": v["stats"]["order_switches"],
                                "survivors": v["stats"]["mixed_survivors"]} for k, v in orders.items()},
    }
    if not args.no_cpu:
        import oracle
        from oracle import queries as OQ
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        ref = OQ.run_query(q, data, oracle.CONV_GPU)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / cpu_s, "unit": "fact rows/s", "cores": threads, "kind": "port",
                               "sample": f"oracle, whole query on all {n} rows, {cpu_s * 1e3:.0f} ms"}
        ok = {k: bool((ref == v["res"]).all()) for k, v in list(runs.items()) + list(orders.items())}
        out["parity"] = all(ok.values())
        out["parity_by_placement"] = ok
    print(json.dumps(out), flush=True)
    eng.close()
    if out.get("parity") is False:
        log("PARITY FAILURE")
        sys.exit(1)


def plain_query_measure(q, data, desc, steps: int, warmup: int, no_cpu: bool = False, cpu_budget: float = 12.0,
                        workload_key: str = ""):
    """One plain per-query step (zero the hash tables and the result, the
    dimension builds, the fused fact call) repeated `steps` times with the
    columns resident in HBM.  The fused call is bracketed with HIP events on
    the stream it runs on (torch's current stream): it holds the compact-table
    pre-pass (when the plan compacts a table) and the fused kernel; the
    kernel alone is in the rocprofv3 summaries under profiles/.  Returns the
    measurement fields of a bench line (value, ms_per_step, config, roofline,
    cpu_baseline, parity)."""
    import torch
    from mordred_amd import abi
    from mordred_amd import query as Q
    abi.load()
    cache = Q.load_cache(data, q)
    gq = Q.GpuQuery(q, data, cache)
    torch.cuda.synchronize()
    log("columns resident in HBM")
    for _ in range(warmup):
        gq.run()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(steps):
        gq.reset()
        gq.build()
        ev[i][0].record(stream)
        gq.probe()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    call_ms = [a.elapsed_time(b) for a, b in ev]
    ms_step = wall * 1e3 / steps
    c_avg = sum(call_ms) / len(call_ms)
    log(f"{q.name}: step {ms_step:.3f} ms, fused call avg {c_avg:.3f} ms (min {min(call_ms):.3f})")
    ncols = len(q.fact_columns())
    alg_bytes = ncols * 4 * data.n
    achieved = alg_bytes / (c_avg * 1e-3) / 1e9
    gpu_res = gq.result()
    out = {
        "value": data.n / (ms_step * 1e-3),
        "ms_per_step": ms_step,
        "config": dict(desc, form="plain per-query GPU (builds + fused fact call per step)"),
        "kernel_rows_per_s": data.n / (c_avg * 1e-3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "frac_kind": "effective (algorithmic bytes)",
                     "traffic": None,
                     "kernel": "pipeline_kernel (fused fact probe+group-by, survivor ring)" if q.kind == "groupby"
                     else "aggr_scan_kernel (fused fact filter+probe+sum, next-tile prefetch)",
                     "algorithmic_bytes": alg_bytes, "call_ms_avg": c_avg,
                     "note": "achieved = algorithmic bytes / fused-call event time (compact-table pre-pass "
                             "+ fused kernel); the kernel alone and its DRAM traffic (rocprofv3 FETCH_SIZE, "
                             "WRITE_SIZE) are in profiles/"},
    }
    t = _profiled_traffic(workload_key, data.n)
    if t:
        out["roofline"]["traffic"] = t["bytes_per_launch"]
        out["roofline"]["traffic_frac"] = t["bytes_per_launch"] / (c_avg * 1e-3) / 1e9 / HBM_PEAK_GBS
        out["roofline"]["traffic_source"] = ("rocprofv3 FETCH_SIZE x2 + WRITE_SIZE of this kernel at this size, "
                                             f"{t['profile']} (measured offline, not in this run)")
    del gq, cache
    torch.cuda.empty_cache()
    if not no_cpu:
        log("oracle: parity check + CPU baseline")
        ref, base = cpu_baseline(q, data, cpu_budget)
        out["cpu_baseline"] = base
        out["parity"] = bool((ref == gpu_res).all())
    else:
        out["cpu_baseline"] = None
        out["parity"] = None
    return out


def _profiled_traffic(key, rows):
    """The workload's DRAM bytes per fused launch from profiles/traffic.json,
    when it was measured at this many rows."""
    try:
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")) as f:
            t = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    return t if t and t.get("rows") == rows else None


def learned_order_phase(eng, q, nseg, steps, learn_runs):
    """Operator order over partly cached segments, chosen by the engine's rule
    (ORDER_AUTO: each order once, then the fastest measured) and by the
    learned plan model (the engine's model, trained further here: EXPLORE
    runs, then USE).  Segments keep only the last join's key and the
    aggregate columns cached (as many as the cache holds), so they run as
    mixed pipelines; the others run on the CPU.  Reports the times, the
    orders and the model's error."""
    from mordred_amd import engine as E
    keep = {q.joins[-1].fact_key} | {a for a in q.aggr if a}
    for s in range(nseg):
        for c in q.fact_columns():
            if c not in keep:
                eng.evict_segment(c, s)
    # (the cache holds half of the segments' columns)
    mixed = min(nseg, int(0.5 * nseg * len(q.fact_columns()) / len(keep)))
    for s in range(mixed):
        for c in keep:
            eng.cache_segment(c, s)
    eng.set_operator_order(E.ORDER_AUTO)
    try:
        rule, rule_orders = [], []
        for _ in range(3 + steps):
            _, st = eng.run(q, E.PLACE_RESIDENCY)
            rule.append(st["total_ms"])
            rule_orders.append(st["operator_order"])
        eng.set_learned_placement(E.LEARNED_EXPLORE)
        for _ in range(learn_runs):
            eng.run(q, E.PLACE_RESIDENCY)
        eng.set_learned_placement(E.LEARNED_USE)
        model, model_orders, by = [], [], {0: 0, 1: 0}
        for _ in range(steps):
            _, st = eng.run(q, E.PLACE_RESIDENCY)
            model.append(st["total_ms"])
            model_orders.append(st["operator_order"])
            by[st["plan_model"]] = by.get(st["plan_model"], 0) + 1
        stats = eng.placement_model_stats()
    finally:
        eng.set_learned_placement(E.LEARNED_OFF)
        eng.set_operator_order(E.ORDER_GPU_FIRST)
    out = {"mixed_segments": mixed, "train_runs": learn_runs,
           "rule_ms_median_after_trials": statistics.median(rule[3:]), "rule_orders": rule_orders,
           "model_ms_median": statistics.median(model), "model_orders": model_orders,
           "plans_by_learned_model": by.get(1, 0), "plans_by_rule": by.get(0, 0), "model": stats}
    log(f"learned order: {out}")
    return out


def batch_roofline(nbytes, gpu_ms, rows, launch):
    """Roofline of the 64-query batch: the effective fraction (algorithmic
    bytes / graph time); and, where profiles/traffic.json holds FETCH_SIZE +
    WRITE_SIZE measured by rocprofv3 for this workload (600M rows, launch
    1024,1,163840), that traffic and its fraction over the same time."""
    out = {"bound": "hbm", "achieved": nbytes / (gpu_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": nbytes / (gpu_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "frac_kind": "effective (algorithmic bytes)",
           "traffic": None, "kernel_ms_avg": gpu_ms,
           "note": "graph = clear + record builds + shared scan + keys; algorithmic bytes = fact columns x 4 B x rows"}
    try:
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")) as f:
            tj = json.load(f)
    except (OSError, ValueError):
        tj = {}
    t = tj.get("batch64")
    if t and rows == 600_000_000 and tuple(launch) == (1024, 1, 163840):
        out["traffic"] = t["bytes_per_launch"]
        out["traffic_frac"] = t["bytes_per_launch"] / (gpu_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
        out["traffic_source"] = ("rocprofv3 FETCH_SIZE x %.2f + WRITE_SIZE x %.2f of this workload, %s (measured "
                                 "offline, not in this run)" % (tj["fetch_scale"], tj["write_scale"], t["profile"]))
    return out


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
    ap.add_argument("--workload", choices=["serve", "q21", "q11", "q4way", "q4way-split", "batch64", "serve-sweep",
                                           "skip", "q21-mixed"], default="serve")
    ap.add_argument("--rate", type=float, default=2000.0, help="serve: offered rate (q/s)")
    ap.add_argument("--rates", default="10,50,200,500,1000,2000", help="serve-sweep: offered rates (q/s)")
    ap.add_argument("--target-chunk-ms", type=float, default=8.0, help="serve: launch-time target")
    ap.add_argument("--rows", type=int, default=None, help="override fact rows")
    ap.add_argument("--learn-runs", type=int, default=48,
                    help="q4way-split: runs that train the learned placement model before it is used")
    ap.add_argument("--batch-config", default=None,
                    help="batch64: threads,rows-per-thread-group,lds-bytes (skips the autotuner)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity check")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1.92 GB device-to-device copy (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU: more GPUs would each scan their own fact "
                                   "segments and all-reduce the int64 group arrays, which is not built here"}))
        sys.exit(2)
    if args.workload == "q4way-split":
        split_bench(args)
        return
    if args.workload == "batch64":
        batch_bench(args)
        return
    if args.workload == "serve":
        serve_headline(args)
        return
    if args.workload == "serve-sweep":
        serve_bench(args)
        return
    if args.workload == "skip":
        skip_bench(args)
        return
    if args.workload == "q21-mixed":
        mixed_bench(args)
        return

    t0 = time.time()
    q, data, desc = workload(args.workload, args.rows)
    log(f"generated {data.n} rows in {time.time() - t0:.1f}s")
    out = plain_query_measure(q, data, desc, args.steps, args.warmup, args.no_cpu, args.cpu_budget, args.workload)
    out = {"metric": METRIC, "value": out.pop("value"), "unit": "fact rows/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": out.pop("ms_per_step"), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32 columns, int64 sums",
           "data": "synthetic, seeded SplitMix64 columns (mordred_amd/datagen.py)", **out}
    print(json.dumps(out), flush=True)
    if out["parity"] is False:
        log("PARITY FAILURE: GPU result differs from the oracle")
        sys.exit(1)


if __name__ == "__main__":
    main()
