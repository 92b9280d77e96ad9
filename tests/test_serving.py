"""Request streams and latency summaries (CPU)."""
from __future__ import annotations

import numpy as np

from mordred_amd import serving


def test_poisson_arrivals():
    a = serving.poisson_arrivals(1000.0, 20_000, seed=1)
    assert a[0] == 0 and np.all(np.diff(a) >= 0)
    rate = len(a) / (a[-1] / 1e3)
    assert abs(rate - 1000.0) / 1000.0 < 0.03
    assert np.array_equal(a, serving.poisson_arrivals(1000.0, 20_000, seed=1))


def test_latency_summary():
    times = [{"arrival_ms": float(i), "done_ms": float(i) + (1.0 if i < 99 else 50.0)} for i in range(100)]
    s = serving.latency_summary(times)
    assert s["p50_ms"] == 1.0 and s["max_ms"] == 50.0
    assert 1.0 < s["p99_ms"] <= 50.0
    assert abs(s["throughput_qps"] - 100 / (149.0 / 1e3)) < 1e-6


def test_request_stream_draws_from_the_mix():
    mixq = ["a", "b", "c"]
    r = serving.request_stream(mixq, 300, seed=2)
    assert set(r) == set(mixq) and len(r) == 300
