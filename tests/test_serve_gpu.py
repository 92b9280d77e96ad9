"""The request loop on the GPU.  Continuous batching (the circular shared
scan) must answer every request with exactly its query's result -- the same
as the query alone, the oracle, and the plain first-come-first-served loop --
whatever launch sizes, wrap-around points and concurrent company it had."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import queries as OQ
from mordred_amd import engine as E
from mordred_amd import mix, serving

pytestmark = pytest.mark.gpu

SEG = oracle.SEGMENT_SIZE
SMALL = {"S": 2000, "P": 30000, "C": 40000}


@pytest.fixture(scope="module")
def sengine(cuda):
    qs = [m.query for m in mix.make_mix(64, 31, SMALL)]
    n = 6 * SEG + 12_345          # 7 segments, the last one partial
    data = mix.make_mix_data(qs, n, seed=31, S=SMALL["S"], P=SMALL["P"], Cn=SMALL["C"])
    eng = E.Engine(device=0, cache_segments=160, threads=8)
    eng.add_mix_data(data, qs)
    refs = {q.name: OQ.run_query(q, data, oracle.CONV_GPU) for q in qs}
    yield qs, data, eng, refs
    eng.close()


def check(reqs, outs, refs):
    for q, r in zip(reqs, outs):
        np.testing.assert_array_equal(r, refs[q.name], err_msg=q.name)


@pytest.mark.parametrize("chunk", [1, 2, 3, 0])
def test_continuous_matches_oracle(sengine, chunk):
    """Launch sizes of 1-3 segments wrap around the 7-segment table at every
    offset; 0 lets the cost model choose."""
    qs, data, eng, refs = sengine
    print(f"[gpu] serve continuous chunk={chunk}", flush=True)
    reqs = serving.request_stream(qs, 150, seed=chunk)
    arr = serving.poisson_arrivals(20_000.0, len(reqs), seed=chunk)
    outs, times, st = eng.serve(reqs, arr, E.SERVE_CONTINUOUS, chunk_segments=chunk, target_chunk_ms=1.0)
    check(reqs, outs, refs)
    assert st["max_active_seen"] > 1
    nseg = 7
    for t in times:
        assert t["arrival_ms"] <= t["admit_ms"] <= t["done_ms"]
        if chunk:
            assert t["launches"] >= -(-nseg // chunk)


def test_each_request_alone_equals_in_the_mix(sengine):
    qs, data, eng, refs = sengine
    reqs = serving.request_stream(qs, 64, seed=9)
    mixed, _, st = eng.serve(reqs, np.zeros(len(reqs)), E.SERVE_CONTINUOUS, chunk_segments=2)
    assert st["max_active_seen"] == 64
    for q, r in zip(reqs[:24], mixed[:24]):
        alone, _, st1 = eng.serve([q], [0.0], E.SERVE_CONTINUOUS, chunk_segments=3)
        assert st1["max_active_seen"] == 1
        np.testing.assert_array_equal(alone[0], r, err_msg=q.name)


def test_plain_loop_matches_oracle(sengine):
    qs, data, eng, refs = sengine
    reqs = serving.request_stream(qs, 40, seed=4)
    outs, times, st = eng.serve(reqs, serving.poisson_arrivals(500.0, 40, seed=4), E.SERVE_PLAIN)
    check(reqs, outs, refs)
    assert st["launches"] == 40 and st["max_active_seen"] == 1


def test_admission_budget_limits_concurrency(sengine):
    """Admission holds the admitted queries' result and touched-bit state
    within the budget (a query is always admitted to an idle device); a 1-byte
    budget serialises the stream.  Every answer stays exact."""
    qs, data, eng, refs = sengine
    reqs = serving.request_stream(qs, 30, seed=5)
    per_query = [24.0 * max(1, q.total_val) + 4.0 * ((max(1, q.total_val) + 31) // 32) for q in reqs]
    budget = 2 * max(per_query)
    outs, times, st = eng.serve(reqs, np.zeros(len(reqs)), E.SERVE_CONTINUOUS, chunk_segments=2,
                                hbm_budget_bytes=budget)
    check(reqs, outs, refs)
    assert st["peak_state_bytes"] <= budget
    outs, times, st = eng.serve(reqs, np.zeros(len(reqs)), E.SERVE_CONTINUOUS, chunk_segments=3,
                                hbm_budget_bytes=1.0)
    check(reqs, outs, refs)
    assert st["max_active_seen"] == 1


def test_learned_model_is_trained_online(sengine):
    qs, data, eng, refs = sengine
    reqs = serving.request_stream(qs, 300, seed=6)
    outs, times, st = eng.serve(reqs, serving.poisson_arrivals(50_000.0, 300, seed=6), E.SERVE_CONTINUOUS,
                                chunk_segments=1, learned=True)
    check(reqs, outs, refs)
    assert st["launches"] >= 20
    assert st["learned_error"] >= 0 and st["analytic_error"] >= 0


def test_lone_requests_run_alone_with_the_same_answers(sengine):
    """lone_plain: a request arriving at an idle loop with nobody waiting runs
    alone through its own fused kernel; sparse arrivals (every request alone)
    and a burst after them (batched) both give the oracle's answers."""
    qs, data, eng, refs = sengine
    reqs = serving.request_stream(qs, 40, seed=12)
    arr = np.concatenate([np.arange(20) * 50.0, np.full(20, 1100.0)])
    print("[gpu] serve lone_plain: 20 sparse requests, then a burst of 20", flush=True)
    outs, times, st = eng.serve(reqs, arr, E.SERVE_CONTINUOUS, chunk_segments=2, lone_plain=True)
    check(reqs, outs, refs)
    assert st["lone_requests"] >= 20
    assert st["max_active_seen"] > 1                       # the burst shared the scan
    for t in times[:20]:
        assert t["launches"] == 1
    # off: every request goes through the shared scan
    outs, times, st = eng.serve(reqs[:5], arr[:5], E.SERVE_CONTINUOUS, chunk_segments=2)
    check(reqs[:5], outs, refs)
    assert st["lone_requests"] == 0


@pytest.mark.parametrize("share", [True, False])
def test_repeated_queries_share_and_stay_exact(sengine, share):
    """A stream of four distinct queries, each repeated many times, arriving
    while the loop runs: identical active queries share their per-launch sums
    (one scans, the copies take its sums), with different starting segments;
    every answer must equal the oracle, with sharing on and off."""
    qs, data, eng, refs = sengine
    big = max(qs, key=lambda q: q.total_val)
    four = [big, qs[1], qs[2], qs[3]]
    reqs = [four[i % 4] for i in range(120)]
    arr = serving.poisson_arrivals(30_000.0, len(reqs), seed=3)
    print(f"[gpu] serve repeated queries, sharing {share}", flush=True)
    try:
        eng.set_batch_sharing(share)
        outs, times, st = eng.serve(reqs, arr, E.SERVE_CONTINUOUS, chunk_segments=2)
    finally:
        eng.set_batch_sharing(True)
    check(reqs, outs, refs)
    assert st["max_active_seen"] > 4
    assert (st["shared_queries"] > 0) == share
