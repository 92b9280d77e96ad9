"""On-disk column files, per-segment min/max and segment skipping (host-only
engine, no GPU).

Skipping must never change a result: every query here is checked against the
oracle over the whole table, with skipping on and off.  The number of skipped
segments is checked against a brute-force count: on a table sorted by
lo_orderdate, a segment's date keys form one run of days and every filter here
keeps one run of days, so min/max excludes a segment exactly when no row of it
passes the filter.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import oracle
from oracle import queries as OQ
from mordred_amd import columnar, datagen
from mordred_amd import engine as E
from mordred_amd import query as Q

SEG = oracle.SEGMENT_SIZE
N = 9 * SEG + 12_345


def brute_minmax(values):
    nseg = (len(values) + SEG - 1) // SEG
    mins = [int(values[s * SEG:(s + 1) * SEG].min()) for s in range(nseg)]
    maxs = [int(values[s * SEG:(s + 1) * SEG].max()) for s in range(nseg)]
    return mins, maxs


def segments_without_survivors(mask):
    nseg = (len(mask) + SEG - 1) // SEG
    return sum(not mask[s * SEG:(s + 1) * SEG].any() for s in range(nseg))


def date_pass(q, dates):
    """Rows whose lo_orderdate passes the query's date filter (by lookup in
    the date table, as the join does)."""
    j = q.join(4)
    table = datagen.make_date()
    col = table[j.filter_col]
    keep = (col >= j.c1) & (col <= j.c2) if j.filter_mode == 1 else (col == j.c1)
    return np.isin(dates, table["d_datekey"][keep])


def host_engine(q, data, threads=4):
    eng = E.Engine(device=-1, threads=threads)
    eng.add_data(data, q)
    return eng


# ---------------------------------------------------------------- file format

def test_file_names_follow_ssb_column_order():
    assert columnar.file_name("lo_orderdate") == "LINEORDER5"
    assert columnar.file_name("lo_orderdate", sorted_table=True) == "LINEORDERSORT5"
    assert columnar.file_name("lo_revenue") == "LINEORDER12"
    assert columnar.file_name("lo_orderkey") == "LINEORDER0"
    assert columnar.file_name("d_year") == "DDATE4"
    assert columnar.file_name("d_yearmonthnum") == "DDATE5"
    assert columnar.file_name("p_brand1") == "PART4"
    assert columnar.file_name("s_region") == "SUPPLIER5"
    assert columnar.file_name("c_nation") == "CUSTOMER4"
    with pytest.raises(KeyError):
        columnar.file_name("lo_nothing")


def test_column_file_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    x = rng.integers(-2**31, 2**31, size=2 * SEG + 999, dtype=np.int64).astype(np.int32)
    eng = E.Engine(device=-1, threads=3)
    eng.add_column("a", x)
    path = tmp_path / "A0"
    eng.store_column_file("a", path)
    # headerless little-endian int32
    assert path.stat().st_size == 4 * len(x)
    np.testing.assert_array_equal(np.fromfile(path, dtype="<i4"), x)
    eng.load_column_file("b", path)                  # whole file
    eng.load_column_file("c", path, n=SEG + 5)       # a prefix, as loadColumn reads num_entries
    assert eng.segments("b") == 3 and eng.segments("c") == 2
    eng.store_column_file("b", tmp_path / "B0")
    eng.store_column_file("c", tmp_path / "C0")
    np.testing.assert_array_equal(np.fromfile(tmp_path / "B0", dtype="<i4"), x)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "C0", dtype="<i4"), x[:SEG + 5])
    with pytest.raises(E.EngineError, match="shorter"):
        eng.load_column_file("d", path, n=len(x) + 1)
    (tmp_path / "odd").write_bytes(b"\x01\x02\x03\x04\x05")
    with pytest.raises(E.EngineError, match="multiple of 4"):
        eng.load_column_file("e", tmp_path / "odd")
    with pytest.raises(E.EngineError, match="cannot open"):
        eng.load_column_file("f", tmp_path / "missing")
    with pytest.raises(E.EngineError, match="duplicate"):
        eng.load_column_file("a", path)
    eng.close()


def test_segment_minmax_is_exact(tmp_path):
    rng = np.random.default_rng(11)
    x = rng.integers(-1000, 1000, size=3 * SEG + 17, dtype=np.int32)
    x[3 * SEG:] = 5000 + np.arange(17)          # partial last segment: its own rows only
    eng = E.Engine(device=-1, threads=2)
    eng.add_column("x", x)
    mins, maxs = eng.segment_minmax("x")
    bmin, bmax = brute_minmax(x)
    assert mins.tolist() == bmin and maxs.tolist() == bmax
    assert (mins[3], maxs[3]) == (5000, 5016)
    # the file format of minmax.cpp:45: "<min> <max>\n" per segment
    eng.store_segment_minmax("x", tmp_path / "xminmax")
    assert (tmp_path / "xminmax").read_text() == "".join(f"{a} {b}\n" for a, b in zip(bmin, bmax))
    eng.close()


def test_minmax_file_validation(tmp_path):
    x = np.arange(2 * SEG + 10, dtype=np.int32)
    eng = E.Engine(device=-1, threads=2)
    eng.add_column("x", x)
    bmin, bmax = brute_minmax(x)

    def load(lines):
        p = tmp_path / "m"
        p.write_text(lines)
        eng.load_segment_minmax("x", p)

    # a looser file covers the data: accepted and used
    load("".join(f"{a - 3} {b + 3}\n" for a, b in zip(bmin, bmax)))
    mins, maxs = eng.segment_minmax("x")
    assert mins.tolist() == [a - 3 for a in bmin] and maxs.tolist() == [b + 3 for b in bmax]
    with pytest.raises(E.EngineError, match="does not cover"):
        load("".join(f"{a + 1} {b}\n" for a, b in zip(bmin, bmax)))
    with pytest.raises(E.EngineError, match="lines for 3 segments"):
        load(f"{bmin[0]} {bmax[0]}\n")
    with pytest.raises(E.EngineError, match="min > max"):
        load("5 4\n0 1\n0 1\n")
    with pytest.raises(E.EngineError, match="expected"):
        load("a b\n")
    # a refused file leaves the last accepted ranges
    assert eng.segment_minmax("x")[0].tolist() == [a - 3 for a in bmin]
    eng.close()


# ---------------------------------------------------------------- skipping

@pytest.fixture(scope="module")
def sorted_q1():
    """A lineorder sorted by lo_orderdate with every column Q1.x reads."""
    q = Q.q11()
    data = Q.make_data(q, N, seed=21, sort_orderdate=True)
    assert np.all(np.diff(data.fact["lo_orderdate"]) >= 0)
    return data


@pytest.mark.parametrize("qname", ["q11", "q12", "q13"])
def test_skipping_sorted_q1x(sorted_q1, qname):
    q = {"q11": Q.q11(), "q12": Q.q12(), "q13": Q.q13()}[qname]
    data = sorted_q1
    ref = OQ.run_query(q, data, oracle.CONV_CPU)
    eng = host_engine(q, data)
    res, st = eng.run(q, E.PLACE_CPU_ONLY)
    np.testing.assert_array_equal(res, ref)
    want = segments_without_survivors(date_pass(q, data.fact["lo_orderdate"]))
    assert want >= 6
    assert st["skipped_segments"] == want
    assert st["cpu_segments"] == 10 - want
    eng.set_skipping(False)
    res2, st2 = eng.run(q, E.PLACE_CPU_ONLY)
    np.testing.assert_array_equal(res2, ref)
    assert st2["skipped_segments"] == 0 and st2["cpu_segments"] == 10
    eng.close()


def test_no_skipping_on_unsorted_table():
    q = Q.q11()
    data = Q.make_data(q, 3 * SEG + 7, seed=4)
    eng = host_engine(q, data)
    res, st = eng.run(q, E.PLACE_CPU_ONLY)
    np.testing.assert_array_equal(res, OQ.run_query(q, data, oracle.CONV_CPU))
    assert st["skipped_segments"] == 0
    eng.close()


def test_fact_filter_skips_segments(sorted_q1):
    """A fact filter on a clustered fact column excludes segments by itself."""
    q = Q.q11()
    data = dataclasses.replace(sorted_q1, fact=dict(sorted_q1.fact))
    data.fact["lo_quantity"] = np.sort(data.fact["lo_quantity"])
    q.joins[0].c1, q.joins[0].c2 = 1992, 1998                 # date filter keeps every row
    eng = host_engine(q, data)
    res, st = eng.run(q, E.PLACE_CPU_ONLY)
    np.testing.assert_array_equal(res, OQ.run_query(q, data, oracle.CONV_CPU))
    qty = data.fact["lo_quantity"]
    want = segments_without_survivors((qty >= 0) & (qty <= 24))
    assert want >= 4 and st["skipped_segments"] == want
    eng.close()


def test_empty_filter_skips_everything(sorted_q1):
    q = Q.q11(year=2001)
    eng = host_engine(q, sorted_q1)
    res, st = eng.run(q, E.PLACE_CPU_ONLY)
    np.testing.assert_array_equal(res, OQ.run_query(q, sorted_q1, oracle.CONV_CPU))
    assert st["skipped_segments"] == 10 and st["cpu_rows"] == 0
    assert not res.any()
    eng.close()


def test_groupby_with_date_filter_skips():
    """Group-by query: only the date join is clustered; the supplier and part
    joins (random keys) exclude nothing."""
    q = Q.q21(2000, 30000)
    q.joins[2].filter_col, q.joins[2].filter_mode, q.joins[2].c1, q.joins[2].c2 = "d_year", 1, 1995, 1996
    data = Q.make_data(q, 6 * SEG + 1000, seed=8, S=2000, P=30000, sort_orderdate=True)
    eng = host_engine(q, data)
    res, st = eng.run(q, E.PLACE_CPU_ONLY)
    np.testing.assert_array_equal(res, OQ.run_query(q, data, oracle.CONV_CPU))
    want = segments_without_survivors(date_pass(q, data.fact["lo_orderdate"]))
    assert want >= 2 and st["skipped_segments"] == want
    eng.close()


def test_load_table_from_files(tmp_path, sorted_q1):
    """Write the sorted table and the date dimension as column + minmax files,
    load them back, and get the same answer and the same skipping."""
    q = Q.q13()
    data = sorted_q1
    columnar.write_table(tmp_path, {c: data.fact[c] for c in q.fact_columns()}, sorted_table=True)
    columnar.write_table(tmp_path, data.tables["date"], sorted_table=False)
    assert (tmp_path / "LINEORDERSORT5").exists() and (tmp_path / "lo_orderdateminmax").exists()
    eng = E.Engine(device=-1, threads=4)
    columnar.load_table(eng, tmp_path, q.fact_columns(), sorted_table=True)
    columnar.load_table(eng, tmp_path, ["d_datekey"])
    res, st = eng.run(q, E.PLACE_CPU_ONLY)
    np.testing.assert_array_equal(res, OQ.run_query(q, data, oracle.CONV_CPU))
    want = segments_without_survivors(date_pass(q, data.fact["lo_orderdate"]))
    assert st["skipped_segments"] == want
    eng.close()
