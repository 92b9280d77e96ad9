"""Request streams and latency statistics for the request loop."""
from __future__ import annotations

import numpy as np


def poisson_arrivals(rate_qps: float, n: int, seed: int) -> np.ndarray:
    """Arrival times (ms, ascending) of n requests of a Poisson process."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1000.0 / rate_qps, n)
    return np.cumsum(gaps) - gaps[0]


def request_stream(mix_queries, n: int, seed: int) -> list:
    """n requests drawn uniformly (with replacement) from a query mix."""
    rng = np.random.default_rng(seed)
    return [mix_queries[int(i)] for i in rng.integers(0, len(mix_queries), n)]


def latency_summary(times) -> dict:
    """p50 / p99 / mean latency (arrival to answer), and throughput over the
    span from the first arrival to the last answer."""
    lat = np.array([t["done_ms"] - t["arrival_ms"] for t in times])
    span_ms = max(t["done_ms"] for t in times) - min(t["arrival_ms"] for t in times)
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "throughput_qps": len(times) / (span_ms / 1e3) if span_ms > 0 else float("inf")}
