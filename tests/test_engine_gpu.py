"""Engine on the GPU: segment cache, CPU/GPU split, PCIe streaming, merge.

Every placement must return the oracle's result buffer bit-exactly: the
cached segment group on the fused GPU kernel, the rest on the engine's CPU
operators and/or streamed to the GPU, merged as CPUProcessing.cu:1786-1819.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import queries as OQ
from mordred_amd import engine as E
from tests import cases

pytestmark = pytest.mark.gpu

SEG = oracle.SEGMENT_SIZE


@pytest.fixture(scope="module", params=["q11", "q21", "q4way"])
def gengine(request, cuda):
    q, data = cases.star(request.param, 5 * SEG + 77_777, seed=9)
    eng = E.Engine(device=0, cache_segments=128, threads=8)
    eng.add_data(data, q)
    ref = OQ.run_query(q, data, oracle.CONV_GPU)
    yield q, data, eng, ref
    eng.close()


@pytest.mark.parametrize("placement", [E.PLACE_RESIDENCY, E.PLACE_COST_SPLIT, E.PLACE_GPU_ONLY, E.PLACE_CPU_ONLY])
def test_placements_match_oracle(gengine, placement):
    q, data, eng, ref = gengine
    print(f"[gpu] engine {q.name} placement {placement}", flush=True)
    eng.place(q, 0.5)
    res, st = eng.run(q, placement)
    np.testing.assert_array_equal(res, ref)
    assert st["gpu_rows"] + st["cpu_rows"] + st["stream_rows"] + st["prefix_rows"] == data.n
    if placement == E.PLACE_RESIDENCY:
        assert st["gpu_segments"] == 3 and st["cpu_segments"] == 3 and st["stream_segments"] == 0
    if placement == E.PLACE_GPU_ONLY:
        assert st["cpu_rows"] == 0 and st["stream_segments"] == 3
    if placement == E.PLACE_CPU_ONLY:
        assert st["gpu_rows"] == 0 and st["stream_rows"] == 0


def test_residency_follows_cost_model(gengine):
    """place(0.5) keeps the full segments with the largest cost weight; the
    partial last segment weighs least and is never the one kept."""
    q, data, eng, ref = gengine
    eng.place(q, 0.5)
    cols = q.fact_columns()
    nseg = eng.segments(cols[0])
    cached = [all(eng.is_cached(c, s) for c in cols) for s in range(nseg)]
    assert sum(cached) == 3 and not cached[nseg - 1]
    eng.place(q, 1.0)
    res, st = eng.run(q, E.PLACE_RESIDENCY)
    np.testing.assert_array_equal(res, ref)
    assert st["gpu_segments"] == nseg and st["cpu_rows"] == 0


def test_partial_last_segment_cached_and_streamed(gengine):
    q, data, eng, ref = gengine
    cols = q.fact_columns()
    nseg = eng.segments(cols[0])
    eng.place(q, 0.0)
    for c in cols:                      # cache only the partial last segment and segment 1
        eng.cache_segment(c, nseg - 1)
        eng.cache_segment(c, 1)
    res, st = eng.run(q, E.PLACE_RESIDENCY)
    np.testing.assert_array_equal(res, ref)
    assert st["gpu_segments"] == 2
    eng.evict_segment(cols[0], 1)       # one column missing: segment 1 leaves the GPU group
    res, st = eng.run(q, E.PLACE_GPU_ONLY)
    np.testing.assert_array_equal(res, ref)
    assert st["gpu_segments"] == 1 and st["stream_segments"] == nseg - 1


# ---------------------------------------------------------------- segment skipping
def _date_filtered_groupby():
    from mordred_amd import query as Q
    q = Q.q21(2000, 30000)
    q.joins[2].filter_col, q.joins[2].filter_mode, q.joins[2].c1, q.joins[2].c2 = "d_year", 1, 1995, 1996
    return q


@pytest.fixture(scope="module", params=["q11", "q21_dated"])
def sorted_engine(request, cuda):
    """A lineorder sorted by lo_orderdate (9 full segments + a partial one)."""
    from mordred_amd import query as Q
    if request.param == "q11":
        q = Q.q11()
        data = Q.make_data(q, 9 * SEG + 12_345, seed=21, sort_orderdate=True)
    else:
        q = _date_filtered_groupby()
        data = Q.make_data(q, 9 * SEG + 12_345, seed=22, S=2000, P=30000, sort_orderdate=True)
    eng = E.Engine(device=0, cache_segments=128, threads=8)
    eng.add_data(data, q)
    ref = OQ.run_query(q, data, oracle.CONV_GPU)
    yield q, data, eng, ref
    eng.close()


@pytest.mark.parametrize("placement", [E.PLACE_RESIDENCY, E.PLACE_COST_SPLIT, E.PLACE_GPU_ONLY])
def test_skipping_on_gpu_placements(sorted_engine, placement):
    """Skipped segments leave every segment group (cached, CPU and streamed);
    place() ranks them last, so the segments the query reads are cached."""
    q, data, eng, ref = sorted_engine
    print(f"[gpu] skipping {q.name} placement {placement}", flush=True)
    eng.set_skipping(True)
    eng.place(q, 0.4)                   # 4 segments: every segment q11 and q21_dated read
    res, st = eng.run(q, placement)
    np.testing.assert_array_equal(res, ref)
    skipped = st["skipped_segments"]
    assert skipped >= 6
    assert (st["gpu_segments"] + st["cpu_segments"] + st["stream_segments"] + st["prefix_segments"]
            == 10 - skipped)
    if placement == E.PLACE_RESIDENCY:
        assert st["cpu_segments"] == 0 and st["gpu_segments"] == 10 - skipped
    eng.set_skipping(False)
    res, st = eng.run(q, placement)
    np.testing.assert_array_equal(res, ref)
    assert st["skipped_segments"] == 0
    eng.set_skipping(True)


# ---------------------------------------------------------------- learned placement
def test_learned_split_plans_are_exact(gengine):
    """COST_SPLIT with the learned placement model: exploration runs random
    candidate plans, and every plan returns the oracle's answer; every run is
    observed by the model."""
    q, data, eng, ref = gengine
    print(f"[gpu] learned split {q.name}", flush=True)
    eng.place(q, 0.5)
    eng.set_learned_placement(E.LEARNED_EXPLORE)
    seen_models, seen_k = set(), set()
    runs = 40
    for _ in range(runs):
        res, st = eng.run(q, E.PLACE_COST_SPLIT)
        np.testing.assert_array_equal(res, ref)
        seen_models.add(st["plan_model"])
        seen_k.add(st["plan_streamed"])
        assert st["plan_streamed"] == st["stream_segments"] + st["prefix_segments"]
        assert st["plan_analytic_ms"] >= 0
    ms = eng.placement_model_stats()
    assert ms["samples"] >= runs and ms["scored"] >= runs - 1
    assert 2 in seen_models and len(seen_k) >= 2
    eng.set_learned_placement(E.LEARNED_USE)
    res, st = eng.run(q, E.PLACE_COST_SPLIT)
    np.testing.assert_array_equal(res, ref)
    assert st["plan_model"] in (0, 1)
    eng.set_learned_placement(E.LEARNED_OFF)


# ---------------------------------------------------------------- mixed pipelines
@pytest.fixture(scope="module", params=["q11", "q21", "q4way"])
def fresh_engine(request, cuda):
    """An engine with nothing of the fact table cached."""
    q, data = cases.star(request.param, 5 * SEG + 77_777, seed=9)
    eng = E.Engine(device=0, cache_segments=256, threads=8)
    eng.add_data(data, q)
    ref = OQ.run_query(q, data, oracle.CONV_GPU)
    yield q, data, eng, ref
    eng.close()


def _uncache_all(eng, q):
    for c in q.fact_columns():
        for s in range(eng.segments(c)):
            eng.evict_segment(c, s)


def test_mixed_pipeline_first_operator_cached(fresh_engine):
    """Only the first operator's fact column is cached (a filter column, or the
    first join's key): the GPU runs that operator over every segment, the CPU
    operators finish the survivors."""
    q, data, eng, ref = fresh_engine
    print(f"[gpu] mixed pipeline {q.name}", flush=True)
    _uncache_all(eng, q)
    col = q.fact_filters[0][0] if q.fact_filters else q.joins[0].fact_key
    nseg = eng.segments(col)
    for s in range(nseg):
        eng.cache_segment(col, s)
    res, st = eng.run(q, E.PLACE_RESIDENCY)
    np.testing.assert_array_equal(res, ref)
    assert st["mixed_segments"] == nseg and st["mixed_rows"] == data.n
    assert st["gpu_segments"] == 0 and st["cpu_segments"] == 0
    assert 0 < st["mixed_survivors"] < data.n


def test_mixed_with_cached_and_plain_segments(fresh_engine):
    """Segments 0-1 fully cached (fused kernel), the key of the last join cached
    in segments 2-3 (mixed), the join keys alternating in the rest (two more
    residency patterns), the partial last segment uncached (CPU operators)."""
    q, data, eng, ref = fresh_engine
    _uncache_all(eng, q)
    cols = q.fact_columns()
    nseg = eng.segments(cols[0])
    for s in (0, 1):
        for c in cols:
            eng.cache_segment(c, s)
    keys = [j.fact_key for j in q.joins]
    for s in (2, 3):
        eng.cache_segment(keys[-1], s)
    for s in range(4, nseg - 1):
        eng.cache_segment(keys[s % len(keys)], s)
    res, st = eng.run(q, E.PLACE_RESIDENCY)
    np.testing.assert_array_equal(res, ref)
    assert st["gpu_segments"] == 2 and st["cpu_segments"] == 1
    assert st["mixed_segments"] == nseg - 3
    res, st = eng.run(q, E.PLACE_COST_SPLIT)
    np.testing.assert_array_equal(res, ref)
    assert st["gpu_segments"] == 2
    assert (st["mixed_segments"] + st["cpu_segments"] + st["stream_segments"] + st["prefix_segments"]
            == nseg - 2)
    res, st = eng.run(q, E.PLACE_GPU_ONLY)
    np.testing.assert_array_equal(res, ref)
    assert st["mixed_segments"] == 0 and st["stream_segments"] == nseg - 2


def test_mixed_filter_then_probe(cuda):
    """A fact filter and a non-date join both cached: the GPU runs filter_GPU2,
    then probe_GPU3 over its row ids; the CPU finishes the part and date
    probes and the group-by."""
    import dataclasses
    q, data = cases.star("q21", 4 * SEG + 4_321, seed=12)
    q = dataclasses.replace(q, fact_filters=[("lo_revenue", 1_000_000, 6_000_000)])
    eng = E.Engine(device=0, cache_segments=64, threads=8)
    eng.add_data(data, q)
    ref = OQ.run_query(q, data, oracle.CONV_GPU)
    nseg = eng.segments("lo_revenue")
    for s in range(nseg):
        eng.cache_segment("lo_revenue", s)
        eng.cache_segment(q.joins[0].fact_key, s)
    print("[gpu] mixed filter then probe", flush=True)
    res, st = eng.run(q, E.PLACE_RESIDENCY)
    np.testing.assert_array_equal(res, ref)
    assert st["mixed_segments"] == nseg and st["gpu_segments"] == 0
    eng.close()


def test_prefix_streaming_cost_split(fresh_engine):
    """COST_SPLIT over an uncached table: when the analytic model prices it
    below whole-segment streaming, the GPU side of the split streams only the
    columns of the operator subset it times fastest, runs those operators,
    and the CPU operators finish the survivors."""
    q, data, eng, ref = fresh_engine
    print(f"[gpu] prefix streaming {q.name}", flush=True)
    _uncache_all(eng, q)
    res, st = eng.run(q, E.PLACE_COST_SPLIT)
    np.testing.assert_array_equal(res, ref)
    nseg = eng.segments(q.fact_columns()[0])
    assert st["gpu_segments"] == 0 and st["mixed_segments"] == 0
    assert st["cpu_segments"] + st["stream_segments"] + st["prefix_segments"] == nseg
    assert st["cpu_rows"] + st["stream_rows"] + st["prefix_rows"] == data.n
    # one plan or the other for the GPU side of the split
    assert st["stream_segments"] == 0 or st["prefix_segments"] == 0
    if len(q.joins) > 1:
        # the part and supplier probes keep a small fraction of the rows (the
        # model picks the subset from its measured per-line and build costs,
        # so which of them it streams depends on the runs before)
        assert 0 < st["est_prefix_ms_per_segment"] < st["est_stream_ms_per_segment"]
        assert st["prefix_segments"] > 0
        assert 0 < st["prefix_survivors"] < st["prefix_rows"] // 2
        # the streamed key columns fit in 24 bits: they crossed packed, in
        # fewer bytes than their 4-byte values
        assert 0 < st["prefix_packed_bytes"] < 4 * st["prefix_rows"] * len(q.fact_columns())
    eng.set_prefix_streaming(False)
    res, st = eng.run(q, E.PLACE_COST_SPLIT)
    eng.set_prefix_streaming(True)
    np.testing.assert_array_equal(res, ref)
    assert st["prefix_segments"] == 0 and st["prefix_packed_bytes"] == 0
    assert st["cpu_segments"] + st["stream_segments"] == nseg


def test_prefix_streaming_with_a_wide_key_column(cuda):
    """A prefix column whose values span more than 2^24 crosses PCIe as 4-byte
    values, beside the columns that cross bit-packed.  lo_partkey is widened
    by multiples of the part table's length: the probe's slot, (key - min) %
    len, is unchanged, so the answer is the original data's (checked on the
    oracle first)."""
    q, data = cases.star("q4way", 5 * SEG + 77_777, seed=9)
    ref = OQ.run_query(q, data, oracle.CONV_GPU)
    j = next(j for j in q.joins if j.fact_key == "lo_partkey")
    rng = np.random.default_rng(5)
    wide = data.fact["lo_partkey"].astype(np.int64) + j.dim_len * rng.integers(0, 1 << 10, data.n)
    assert wide.max() - wide.min() >= 1 << 24 and wide.max() < 2**31
    data.fact["lo_partkey"] = wide.astype(np.int32)
    np.testing.assert_array_equal(OQ.run_query(q, data, oracle.CONV_GPU), ref)
    print("[gpu] prefix streaming with a wide lo_partkey", flush=True)
    eng = E.Engine(device=0, cache_segments=256, threads=8)
    try:
        eng.add_data(data, q)
        _uncache_all(eng, q)
        for _ in range(3):  # later runs plan with the measured CPU costs
            res, st = eng.run(q, E.PLACE_COST_SPLIT)
            np.testing.assert_array_equal(res, ref)
            print(f"[gpu]   prefix {st['prefix_segments']} segments, {st['prefix_packed_bytes']} packed bytes",
                  flush=True)
            assert st["cpu_rows"] + st["stream_rows"] + st["prefix_rows"] == data.n
            # the packed columns (supplier and customer keys, date) take
            # fewer than 4 bytes per row each
            assert st["prefix_packed_bytes"] < 4 * st["prefix_rows"] * (len(q.fact_columns()) - 1) + 1
    finally:
        eng.close()


@pytest.mark.parametrize("order", [E.ORDER_GPU_FIRST, E.ORDER_CPU_FIRST, E.ORDER_ALTERNATE])
def test_operator_orders(fresh_engine, order):
    """The reference's operator orders over partly cached segments: the
    aggregate columns and the last join's key cached everywhere (q11's
    lo_discount is also its first fact filter's column), the other keys and
    filter columns on the host.  CPU-first runs the CPU filters, the
    GPU filters and probes, the CPU probes, then the group-by on the GPU;
    alternate runs the GPU filters, the CPU filters and probes, then the GPU
    probes and group-by.  Every order returns the oracle's answer."""
    q, data, eng, ref = fresh_engine
    print(f"[gpu] operator order {order} {q.name}", flush=True)
    _uncache_all(eng, q)
    nseg = eng.segments(q.fact_columns()[0])
    cached = {q.joins[-1].fact_key} | {a for a in q.aggr if a}
    for c in cached:
        for s in range(nseg):
            eng.cache_segment(c, s)
    eng.set_operator_order(order)
    try:
        res, st = eng.run(q, E.PLACE_RESIDENCY)
    finally:
        eng.set_operator_order(E.ORDER_GPU_FIRST)
    np.testing.assert_array_equal(res, ref)
    assert st["operator_order"] == order
    assert st["mixed_segments"] == nseg and st["gpu_segments"] == 0 and st["cpu_segments"] == 0
    assert 0 < st["mixed_survivors"] <= data.n
    if order == E.ORDER_GPU_FIRST:
        assert st["order_switches"] == 0
    else:
        # at least one switch per chunk of 16 segments; with a cached fact
        # filter (q11: lo_discount) the alternate order switches back
        chunks = (nseg + 15) // 16
        back = order == E.ORDER_ALTERNATE and q.fact_filters
        assert st["order_switches"] >= (2 if back else 1) * chunks


def test_operator_order_auto(fresh_engine):
    """MORDRED_ORDER_AUTO (main.cu:155-161 keeps the faster fact pipeline):
    the first three runs of a query and residency try GPU-first, CPU-first
    and alternate once each; the fourth uses one of them; all are exact."""
    q, data, eng, ref = fresh_engine
    print(f"[gpu] operator order auto {q.name}", flush=True)
    _uncache_all(eng, q)
    nseg = eng.segments(q.fact_columns()[0])
    for c in {q.joins[-1].fact_key} | {a for a in q.aggr if a}:
        for s in range(nseg):
            eng.cache_segment(c, s)
    eng.set_operator_order(E.ORDER_AUTO)
    try:
        seen = []
        for _ in range(4):
            res, st = eng.run(q, E.PLACE_RESIDENCY)
            np.testing.assert_array_equal(res, ref)
            assert st["mixed_segments"] == nseg
            seen.append(st["operator_order"])
    finally:
        eng.set_operator_order(E.ORDER_GPU_FIRST)
    assert seen[:3] == [E.ORDER_GPU_FIRST, E.ORDER_CPU_FIRST, E.ORDER_ALTERNATE]
    assert seen[3] in seen[:3]


def test_learned_operator_order_and_residency_are_exact(fresh_engine):
    """The learned plan model also picks the operator order of partly cached
    segments (engine order AUTO) and weighs the segments place() keeps.
    Exploration runs random orders; FORCE lets the model choose once it has
    been trained; every run, and the runs after a model-weighted place(),
    return the oracle's answer."""
    q, data, eng, ref = fresh_engine
    print(f"[gpu] learned order + residency {q.name}", flush=True)
    _uncache_all(eng, q)
    nseg = eng.segments(q.fact_columns()[0])
    for c in {q.joins[-1].fact_key} | {a for a in q.aggr if a}:
        for s in range(nseg):
            eng.cache_segment(c, s)
    eng.set_operator_order(E.ORDER_AUTO)
    try:
        eng.set_learned_placement(E.LEARNED_EXPLORE)
        orders = set()
        for _ in range(24):
            res, st = eng.run(q, E.PLACE_RESIDENCY)
            np.testing.assert_array_equal(res, ref)
            assert st["mixed_segments"] == nseg
            orders.add(st["operator_order"])
        assert len(orders) >= 2
        before = eng.placement_model_stats()
        assert before["samples"] >= 24
        eng.set_learned_placement(E.LEARNED_FORCE)
        for _ in range(4):
            res, st = eng.run(q, E.PLACE_RESIDENCY)
            np.testing.assert_array_equal(res, ref)
            assert st["plan_model"] == 1 and st["plan_predicted_ms"] > 0
        eng.place(q, 0.5)
        after = eng.placement_model_stats()
        assert after["residency_calls"] == before["residency_calls"] + 1
        for placement in (E.PLACE_COST_SPLIT, E.PLACE_RESIDENCY):
            res, st = eng.run(q, placement)
            np.testing.assert_array_equal(res, ref)
    finally:
        eng.set_learned_placement(E.LEARNED_OFF)
        eng.set_operator_order(E.ORDER_GPU_FIRST)


def test_aggregate_query_with_other_join_rejected_on_gpu(cuda):
    """An aggregate query that joins a table other than the date table is
    refused under every placement (the aggregate operators probe only table 4,
    so placements would disagree); nothing is launched for it."""
    import copy
    from mordred_amd import datagen
    from mordred_amd import query as Q
    q = Q.q11(1994)
    data = Q.make_data(q, SEG + 999, seed=8)
    data.tables["supplier"] = datagen.make_supplier(500, 8)
    data.fact["lo_suppkey"] = (np.arange(data.n) % 500 + 1).astype(np.int32)
    bad = copy.deepcopy(q)
    bad.joins = [Q.DimJoin(1, "supplier", "lo_suppkey", "s_suppkey", 500, 0, filter_col="s_region", c1=1,
                           c2=1)] + bad.joins
    eng = E.Engine(device=0, cache_segments=16, threads=2)
    eng.add_data(data, bad)
    for placement in (E.PLACE_RESIDENCY, E.PLACE_COST_SPLIT, E.PLACE_GPU_ONLY, E.PLACE_CPU_ONLY):
        with pytest.raises(E.EngineError, match="aggregate queries join only table 4"):
            eng.run(bad, placement)
    res, _ = eng.run(q, E.PLACE_GPU_ONLY)
    assert oracle.res_scalar(res) == oracle.res_scalar(OQ.run_query(q, data, oracle.CONV_GPU))
    eng.close()
