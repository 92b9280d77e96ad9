"""Shared-scan batches on the GPU: 64 mix queries in one persistent kernel.

Every query's result must equal the oracle (GPU convention) and its own plain
engine run bit for bit, alone or in the mix, under every launch configuration
the autotuner tries, and across graph replays with different batches.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import oracle
from oracle import queries as OQ
from mordred_amd import engine as E
from mordred_amd import mix

pytestmark = pytest.mark.gpu

SEG = oracle.SEGMENT_SIZE
SMALL = {"S": 2000, "P": 30000, "C": 40000}


@pytest.fixture(scope="module")
def bengine(cuda):
    qs = [m.query for m in mix.make_mix(64, 11, SMALL)]
    other = [m.query for m in mix.make_mix(64, 12, SMALL)]
    n = 2 * SEG + 77_777
    data = mix.make_mix_data(qs + other, n, seed=11, S=SMALL["S"], P=SMALL["P"], Cn=SMALL["C"])
    eng = E.Engine(device=0, cache_segments=96, threads=8)
    eng.add_mix_data(data, qs + other)
    refs = {q.name: OQ.run_query(q, data, oracle.CONV_GPU) for q in qs + other}
    yield qs, other, data, eng, refs
    eng.close()


def check(qs, outs, refs):
    for q, r in zip(qs, outs):
        np.testing.assert_array_equal(r, refs[q.name], err_msg=q.name)


def test_batch_of_64_matches_oracle(bengine):
    qs, _, _, eng, refs = bengine
    print("[gpu] batch of 64, default config", flush=True)
    outs, st = eng.batch(qs)
    check(qs, outs, refs)
    assert st["lds_queries"] + st["global_queries"] == 64


@pytest.mark.parametrize("config", [(256, 1, 16 << 10), (512, 2, 96 << 10), (1024, 1, 160 << 10),
                                    (256, 2, 12 << 10)])
def test_batch_configs_match_oracle(bengine, config):
    """LDS budgets from nearly descriptors-only (group arrays in HBM) to 160 KiB."""
    qs, _, _, eng, refs = bengine
    print(f"[gpu] batch config {config}", flush=True)
    outs, st = eng.batch(qs, config)
    check(qs, outs, refs)
    if config[2] == 12 << 10:
        assert st["global_queries"] >= 32


@pytest.mark.parametrize("forms", [E.Engine.FORM_DICT | E.Engine.FORM_ATTR, E.Engine.FORM_ATTR,
                                   E.Engine.FORM_MASK, E.Engine.FORM_ALL])
def test_dimension_forms_match_oracle(bengine, forms):
    """Dictionary codes, attribute entries (here on every table) and masks."""
    qs, _, _, eng, refs = bengine
    print(f"[gpu] batch forms {forms}", flush=True)
    try:
        eng.set_batch_forms(forms)
        outs, st = eng.batch(qs, (1024, 1, 160 << 10))
        check(qs, outs, refs)
        outs, st = eng.batch(qs[:9], (512, 1, 64 << 10))
        check(qs[:9], outs, refs)
    finally:
        eng.set_batch_forms(E.Engine.FORM_ALL)


def test_each_query_alone_equals_mix(bengine):
    qs, _, _, eng, refs = bengine
    mixed, _ = eng.batch(qs)
    for i, q in enumerate(qs):
        alone, _ = eng.batch([q])
        np.testing.assert_array_equal(alone[0], mixed[i], err_msg=q.name)


def test_batch_equals_plain_engine_runs(bengine):
    qs, _, _, eng, refs = bengine
    outs, _ = eng.batch(qs[:16])
    for q, r in zip(qs[:16], outs):
        plain, _ = eng.run(q, E.PLACE_RESIDENCY)
        np.testing.assert_array_equal(r, plain, err_msg=q.name)


def test_graph_replays_across_batches(bengine):
    """One captured graph serves batches with different queries: only the
    device descriptor changes between them."""
    qs, other, _, eng, refs = bengine
    cfg = (512, 1, 64 << 10)
    a1, s1 = eng.batch(qs, cfg)
    b1, s2 = eng.batch(other, cfg)
    a2, s3 = eng.batch(qs, cfg)
    assert s2["graph_captured"] == 0 and s3["graph_captured"] == 0
    check(qs, a1, refs)
    check(other, b1, refs)
    check(qs, a2, refs)


def test_autotuner_configs_agree(bengine):
    qs, _, _, eng, refs = bengine
    print("[gpu] batch autotune", flush=True)
    best, table = eng.batch_tune(qs, reps=1)
    assert best in [c for c, ms in table if ms >= 0]
    assert sum(ms >= 0 for _, ms in table) >= 8
    outs, st = eng.batch(qs)
    assert st["config"] == best
    check(qs, outs, refs)


def test_wide_attribute_words(bengine):
    """A query grouping on d_datekey (16 bits of the date table's attribute
    word, with d_year and d_yearmonthnum beside it): the batch keeps its
    attribute words unpacked, one per dimension, in the pair loop."""
    qs, _, data, eng, refs = bengine
    from mordred_amd.query import DimJoin, StarQuery
    from mordred_amd import datagen
    span = int(data.tables["date"]["d_datekey"].max()) - datagen.DATE_MIN_KEY + 1
    date = DimJoin(4, "date", "lo_orderdate", "d_datekey", datagen.DATE_HT_LEN, datagen.DATE_MIN_KEY,
                   filter_col="d_year", c1=1994, c2=1995, group_col="d_datekey",
                   min_val=datagen.DATE_MIN_KEY, unique_val=1)
    supp = DimJoin(1, "supplier", "lo_suppkey", "s_suppkey", SMALL["S"], 0, filter_col="s_region", c1=2, c2=2)
    wide = StarQuery("wide_datekey", [supp, date], [], ("lo_revenue", None), "groupby", span)
    batch = [wide] + [q for q in qs if q.join(4) is not None][:8]
    print(f"[gpu] batch with a d_datekey group ({span} groups)", flush=True)
    outs, st = eng.batch(batch, (1024, 1, 160 << 10))
    np.testing.assert_array_equal(outs[0], OQ.run_query(wide, data, oracle.CONV_GPU), err_msg=wide.name)
    assert (outs[0][4::6] != 0).any()
    check(batch[1:], outs[1:], refs)


def test_identical_queries_share_sums(bengine):
    """Repeated queries in one batch: every copy after the first takes the
    first's sums (when those sit in LDS) instead of scanning; every copy's
    result must still equal the oracle, with sharing on and off."""
    qs, _, _, eng, refs = bengine
    big = max(qs, key=lambda q: q.total_val)
    batch = [qs[3], big, qs[5], qs[3], qs[7], big, qs[3], qs[5]] + qs[10:20]
    print(f"[gpu] batch with repeated queries (largest {big.name}, {big.total_val} groups)", flush=True)
    try:
        outs, st = eng.batch(batch, (1024, 1, 160 << 10))
        check(batch, outs, refs)
        assert st["shared_queries"] >= 3
        eng.set_batch_sharing(False)
        outs2, st2 = eng.batch(batch, (1024, 1, 160 << 10))
        assert st2["shared_queries"] == 0
        check(batch, outs2, refs)
    finally:
        eng.set_batch_sharing(True)


def test_dead_query_and_rejections(bengine):
    qs, _, _, eng, refs = bengine
    q = dataclasses.replace(qs[0], fact_filters=[("lo_revenue", 900_000, 100_000)])
    outs, st = eng.batch([q, qs[1]])
    assert st["dead_queries"] == 1 and not outs[0].any()
    np.testing.assert_array_equal(outs[1], refs[qs[1].name])
    with pytest.raises(E.EngineError, match="1..64"):
        eng.batch(qs + qs[:1])


def test_hash_collision_is_reported(bengine):
    """A dimension table shorter than its key range maps two rows to one slot:
    the record tables cannot stand for it, and the batch says so."""
    qs, _, _, eng, _ = bengine
    q = next(q for q in qs if any(j.slot == 1 for j in q.joins))
    joins = [dataclasses.replace(j, dim_len=j.dim_len // 2) if j.slot == 1 else j for j in q.joins]
    bad = dataclasses.replace(q, joins=joins)
    with pytest.raises(E.EngineError, match="collision"):
        eng.batch([bad])


def test_empty_dimension_range_is_dead(bengine):
    """A dimension filter BETWEEN c1 AND c2 with c1 > c2 keeps no dimension
    row (build_GPU2 mode 1): on a key or wide column the batch marks the query
    dead and returns its empty result; on a narrow attribute the mask table
    rejects every row."""
    qs, _, data, eng, refs = bengine
    q = next(q for q in qs if any(j.filter_col and j.filter_col == j.dim_key for j in q.joins))
    joins = [dataclasses.replace(j, c1=j.c2 + 1, c2=j.c2) if j.filter_col == j.dim_key else j for j in q.joins]
    bad = dataclasses.replace(q, joins=joins)
    print(f"[gpu] empty key range {q.name}", flush=True)
    outs, st = eng.batch([bad, qs[1]])
    assert st["dead_queries"] == 1 and not outs[0].any()
    np.testing.assert_array_equal(outs[1], refs[qs[1].name])
    np.testing.assert_array_equal(outs[0], OQ.run_query(bad, data, oracle.CONV_GPU))


def test_interval_tables_match_range_compares(bengine):
    """Many key-range filters on one dimension share an interval table when
    four or more queries filter the same column; with few queries they stay
    range compares.  Both give each query's own answer."""
    qs, other, _, eng, refs = bengine
    print("[gpu] interval tables", flush=True)
    keyed = [q for q in qs + other if any(j.filter_col and j.filter_col == j.dim_key for j in q.joins)]
    assert len(keyed) >= 8
    for sub in (keyed[:3], keyed[:12], keyed):
        outs, _ = eng.batch(sub[:64])
        check(sub[:64], outs, refs)
