"""The engine's CPU half (host-only engine, no GPU) against the oracle.

The engine's CPU operators are product code, separate from the oracle; they
must give the oracle's answers bit-exactly (CPU convention: only non-zero key
words are written, CPUProcessing.cu:498-501).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import queries as OQ
from mordred_amd import engine as E
from tests import cases, semantics

SEG = oracle.SEGMENT_SIZE


@pytest.fixture(scope="module", params=["q11", "q21", "q4way"])
def host_engine(request):
    q, data = cases.star(request.param, 2 * SEG + 54_321, seed=5)
    eng = E.Engine(device=-1, threads=4)
    eng.add_data(data, q)
    yield q, data, eng
    eng.close()


def test_cpu_only_matches_oracle(host_engine):
    q, data, eng = host_engine
    res, st = eng.run(q, E.PLACE_CPU_ONLY)
    ref = OQ.run_query(q, data, oracle.CONV_CPU)
    np.testing.assert_array_equal(res, ref)
    assert st["cpu_rows"] == data.n and st["gpu_rows"] == 0
    if q.kind == "groupby":
        assert semantics.decode_groups(res) == semantics.group_map(q, data)


@pytest.mark.parametrize("name", cases.golden_names())
def test_cpu_only_golden(name):
    q, data, exp = cases.load_golden(name)
    eng = E.Engine(device=-1, threads=3)
    eng.add_data(data, q)
    res, _ = eng.run(q, E.PLACE_CPU_ONLY)
    if q.kind == "aggr":
        assert oracle.res_scalar(res) == exp["scalar"]
    else:
        assert semantics.decode_groups(res) == exp["groups"]
    eng.close()


def test_host_engine_rejects_gpu_placements(host_engine):
    q, data, eng = host_engine
    with pytest.raises(E.EngineError):
        eng.run(q, E.PLACE_RESIDENCY)
    with pytest.raises(E.EngineError):
        eng.cache_segment(q.fact_columns()[0], 0)


def test_bad_queries_are_rejected():
    eng = E.Engine(device=-1, threads=1)
    eng.add_column("a", np.arange(10, dtype=np.int32))
    from mordred_amd import query as Q
    q = Q.StarQuery("bad", [], [], ("missing", None), "aggr", 1)
    with pytest.raises(E.EngineError, match="unknown aggregate"):
        eng.run(q, E.PLACE_CPU_ONLY)
    with pytest.raises(E.EngineError, match="duplicate"):
        eng.add_column("a", np.arange(3, dtype=np.int32))
    eng.close()


def test_learned_placement_needs_a_gpu(host_engine):
    q, data, eng = host_engine
    eng.set_learned_placement(E.LEARNED_OFF)
    with pytest.raises(E.EngineError, match="needs a GPU"):
        eng.set_learned_placement(E.LEARNED_USE)
    with pytest.raises(E.EngineError, match="bad learned placement mode"):
        eng.set_learned_placement(7)
    assert eng.placement_model_stats()["samples"] == 0


def test_binding_struct_sizes_match_library():
    """The ctypes copies of the engine header's structs have the library's
    sizes (a field added out of order on one side would let the library
    write past the binding's buffer)."""
    lib = E._lib()
    pairs = {
        "mordred_join": E.Join, "mordred_fact_filter": E.FactFilter, "mordred_star_query": E.StarQueryC,
        "mordred_run_stats": E.RunStats, "mordred_placement_stats": E.PlacementStats,
        "mordred_batch_config": E.BatchConfig, "mordred_batch_stats": E.BatchStats,
        "mordred_batch_tune_entry": E.BatchTuneEntry, "mordred_serve_opts": E.ServeOpts,
        "mordred_serve_times": E.ServeTimes, "mordred_serve_stats": E.ServeStats,
    }
    import ctypes
    for name, cls in pairs.items():
        assert lib.mordred_engine_struct_size(name.encode()) == ctypes.sizeof(cls), name
    assert lib.mordred_engine_struct_size(b"no_such_struct") == 0


def test_aggregate_query_joins_only_the_date_table():
    """The aggregate operators probe only table 4 (filter_probe_aggr_GPU2,
    filter_probe_aggr_CPU), so an aggregate query with another join would give
    placement-dependent answers; the engine refuses it (resolve() runs before
    any placement; tests/test_engine_gpu.py covers the GPU placements).  The
    date-only query still equals the oracle."""
    import copy
    from mordred_amd import query as Q
    q = Q.q11(1994)
    data = Q.make_data(q, SEG + 999, seed=8)
    bad = copy.deepcopy(q)
    data.tables["supplier"] = __import__("mordred_amd").datagen.make_supplier(500, 8)
    data.fact["lo_suppkey"] = (np.arange(data.n) % 500 + 1).astype(np.int32)
    bad.joins = [Q.DimJoin(1, "supplier", "lo_suppkey", "s_suppkey", 500, 0, filter_col="s_region", c1=1, c2=1)] + bad.joins
    eng = E.Engine(device=-1, threads=2)
    eng.add_data(data, bad)
    with pytest.raises(E.EngineError, match="aggregate queries join only table 4"):
        eng.run(bad, E.PLACE_CPU_ONLY)
    res, _ = eng.run(q, E.PLACE_CPU_ONLY)
    assert oracle.res_scalar(res) == oracle.res_scalar(OQ.run_query(q, data, oracle.CONV_CPU))
    eng.close()
