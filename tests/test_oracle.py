"""Pins the oracle (C restatement of src/gpudb/CPUProcessing.cu).

Checks, each chosen so that a plausible slip in the oracle fails at least one:
  * golden fixtures from a pure-Python brute force (tests/golden);
  * the pandas statement of the same SQL (tests/semantics.py) on seeded
    multi-segment inputs with a permuted segment group and a partial last
    segment;
  * invariants: sum over groups == ungrouped sum, offset lists are the
    qualifying rows, CPU/GPU conventions agree on the valid domain;
  * closed forms and special cases: empty input, nothing / everything
    selected, keys at the ends of the hash range, group code 0.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import (CONV_CPU, CONV_GPU, BuildArgs, FilterArgs, GroupbyArgs, Offset, ProbeArgs, iptr,
                    lib, sptr)
from oracle import queries as OQ
from tests import cases, semantics

SEG = oracle.SEGMENT_SIZE


@pytest.fixture(scope="module", autouse=True)
def _threads():
    oracle.set_threads(4)
    yield


# ------------------------------------------------------------- golden vectors

@pytest.mark.parametrize("name", cases.golden_names())
@pytest.mark.parametrize("conv", [CONV_CPU, CONV_GPU])
def test_golden_query(name, conv):
    q, data, exp = cases.load_golden(name)
    res = OQ.run_query(q, data, conv)
    if q.kind == "aggr":
        assert oracle.res_scalar(res) == exp["scalar"]
    else:
        got = semantics.decode_groups(res)
        if conv == CONV_CPU:
            # CPU path writes only non-zero keys (CPUProcessing.cu:498-501); the
            # zero keys are zero anyway because results start zeroed
            pass
        assert got == exp["groups"]


@pytest.mark.parametrize("name", cases.golden_names())
def test_golden_hash_tables(name):
    q, data, exp = cases.load_golden(name)
    hts = OQ.build_tables(q, data, CONV_CPU)
    for slot, ht in hts.items():
        np.testing.assert_array_equal(ht, exp["ht"][slot])
    hts_gpu = OQ.build_tables(q, data, CONV_GPU)
    for slot, ht in hts_gpu.items():
        np.testing.assert_array_equal(ht, exp["ht"][slot])


@pytest.mark.parametrize("name", cases.golden_names())
def test_golden_probe_offsets(name):
    """probe_CPU over all joined tables yields exactly the brute-force
    (lo_row, dim rows) of the rows that pass the joins (fact filters are
    applied by filtering the brute-force list)."""
    q, data, exp = cases.load_golden(name)
    hts = OQ.build_tables(q, data, CONV_CPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    n = data.n
    out = [np.zeros(n, dtype=np.int32) for _ in range(5)]
    total = np.zeros(1, dtype=np.int32)
    lib.oracle_probe_CPU(p, Offset(*[iptr(a) for a in out]), n, iptr(total), 0,
                         sptr(OQ.segment_group(n)), CONV_CPU)
    m = int(total[0])
    got = np.stack([a[:m] for a in out], axis=1)
    got = got[np.lexsort(got.T[::-1])]
    if q.fact_filters:
        keep_rows = np.ones(m, dtype=bool)
        for c, lo, hi in q.fact_filters:
            x = data.fact[c][got[:, 0]]
            keep_rows &= (x >= lo) & (x <= hi)
        got = got[keep_rows]
    np.testing.assert_array_equal(got, exp["probe_rows"])


# ------------------------------------------------------- pandas statement

@pytest.fixture(scope="module", params=["q11", "q21", "q4way"])
def star_case(request):
    q, data = cases.star(request.param, 2 * SEG + 123_457, seed=97)
    return q, data


def test_oracle_matches_pandas_permuted_segments(star_case):
    q, data = star_case
    # segment group [1, 0, 2]: the partial last segment stays last
    res = OQ.run_query(q, data, CONV_CPU, segments=[1, 0, 2], n=data.n)
    if q.kind == "aggr":
        assert oracle.res_scalar(res) == semantics.scalar_sum(q, data)
    else:
        assert semantics.decode_groups(res) == semantics.group_map(q, data)


def test_oracle_segment_subset(star_case):
    """A segment group that holds only segment 1 covers exactly rows 2^20..2^21."""
    q, data = star_case
    res = OQ.run_query(q, data, CONV_GPU, segments=[1], n=SEG)
    rows = np.arange(SEG, 2 * SEG)
    if q.kind == "aggr":
        assert oracle.res_scalar(res) == semantics.scalar_sum(q, data, rows)
    else:
        assert semantics.decode_groups(res) == semantics.group_map(q, data, rows)


def test_group_sum_equals_ungrouped_sum(star_case):
    q, data = star_case
    if q.kind != "groupby":
        pytest.skip("group-by invariant")
    res = OQ.run_query(q, data, CONV_GPU)
    df = semantics.joined(q, data)
    assert int(oracle.res_sums(res).sum()) == int((df["a1"] - df["a2"]).sum())


def test_conventions_agree_on_valid_domain(star_case):
    q, data = star_case
    a = OQ.run_query(q, data, CONV_CPU)
    b = OQ.run_query(q, data, CONV_GPU)
    assert semantics.decode_groups(a) == semantics.decode_groups(b)
    np.testing.assert_array_equal(oracle.res_sums(a), oracle.res_sums(b))


def test_dense_index_layout_known_answer():
    """Q2.1 strides (QueryOptimizer.cu:2909-2919): group (brand b, year y) sits
    at h = 7*b + (y-1992); keys are (supp value 0, 0, brand, year)."""
    q, data, exp = cases.load_golden("q21_tiny")
    res = OQ.run_query(q, data, CONV_GPU).reshape(-1, 6)
    for (v1, v2, brand, year), s in exp["groups"].items():
        h = 7 * brand + (year - 1992)
        assert tuple(res[h, :4]) == (v1, v2, brand, year)
        assert res[h, 4:6].copy().view(np.int64)[0] == s


# ------------------------------------------------------------ compaction

def _filter_args(col1=None, c=(0, 0, 0, 0), modes=(1, 1), col2=None):
    f = FilterArgs()
    f.filter_col1, f.filter_col2 = iptr(col1), iptr(col2)
    f.compare1, f.compare2, f.compare3, f.compare4 = c
    f.mode1, f.mode2 = modes
    return f


@pytest.mark.parametrize("modes", [(0, 0), (1, 1), (2, 2), (1, 2), (2, 0), (7, 1)])
@pytest.mark.parametrize("conv", [CONV_CPU, CONV_GPU])
def test_filter_modes(modes, conv):
    rng = np.random.default_rng(3)
    n = SEG + 5000
    a = rng.integers(0, 10, n).astype(np.int32)
    b = rng.integers(0, 10, n).astype(np.int32)
    c = (2, 5, 3, 7)
    out = np.zeros(n, dtype=np.int32)
    total = np.zeros(1, dtype=np.int32)
    lib.oracle_filter_CPU(_filter_args(a, c, modes, b), iptr(out), n, iptr(total), 0,
                          sptr(OQ.segment_group(n)), conv)
    got = np.sort(out[: total[0]])

    def mode1(x):
        m = modes[0]
        if m == 0:
            return x == c[0] if conv == CONV_GPU else np.ones_like(x, dtype=bool)
        if m == 1:
            return (x >= c[0]) & (x <= c[1])
        if m == 2:
            return (x == c[0]) | (x == c[1])
        return np.ones_like(x, dtype=bool)

    f1 = mode1(a)
    m2 = modes[1]
    if m2 == 0:
        sel = (f1 & (b == c[2])) if conv == CONV_GPU else f1
    elif m2 == 1:
        sel = f1 & (b >= c[2]) & (b <= c[3])
    elif m2 == 2:
        sel = ((f1 & (b == c[2])) | (b == c[3])) if conv == CONV_GPU else (f1 & ((b == c[2]) | (b == c[3])))
    else:
        sel = f1
    np.testing.assert_array_equal(got, np.nonzero(sel)[0].astype(np.int32))


def test_filter_offset_mode_and_filter_probe_pipeline(star_case):
    """filter_CPU2 over an offset list, then filter_probe_CPU, then
    aggregationCPU reproduce the fused Q1.1 result."""
    q, data = star_case
    if q.kind != "aggr":
        pytest.skip("Q1.1 pipeline")
    hts = OQ.build_tables(q, data, CONV_CPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    n = data.n
    sg = OQ.segment_group(n)
    # step 1: filter1 only (discount) -> offsets
    f1 = _filter_args(data.fact["lo_discount"], (1, 3, 0, 0), (1, 1))
    off1 = np.zeros(n, dtype=np.int32)
    t1 = np.zeros(1, dtype=np.int32)
    lib.oracle_filter_CPU(f1, iptr(off1), n, iptr(t1), 0, sptr(sg), CONV_CPU)
    m1 = int(t1[0])
    # step 2: filter2 over the offset list (quantity)
    f2 = _filter_args(data.fact["lo_quantity"], (0, 24, 0, 0), (1, 1))
    off2 = np.zeros(m1, dtype=np.int32)
    t2 = np.zeros(1, dtype=np.int32)
    lib.oracle_filter_CPU2(iptr(off1), f2, iptr(off2), m1, iptr(t2), 0, CONV_CPU)
    m2 = int(t2[0])
    # step 3: probe the date table from the offset list
    lo3, d3 = np.zeros(m2, dtype=np.int32), np.zeros(m2, dtype=np.int32)
    t3 = np.zeros(1, dtype=np.int32)
    p4 = ProbeArgs()
    p4.key_col4, p4.ht4, p4.dim_len4, p4.min_key4 = p.key_col4, p.ht4, p.dim_len4, p.min_key4
    lib.oracle_probe_CPU2(Offset(iptr(off2), None, None, None, None), p4,
                          Offset(iptr(lo3), None, None, None, iptr(d3)), m2, iptr(t3), 0, CONV_CPU)
    m3 = int(t3[0])
    res = np.zeros(6, dtype=np.int32)
    lib.oracle_aggregationCPU(iptr(lo3), g, m3, iptr(res))
    assert oracle.res_scalar(res) == semantics.scalar_sum(q, data)
    # the dimension rows are those of the date keys
    dk = data.tables["date"]["d_datekey"]
    np.testing.assert_array_equal(dk[d3[:m3]], data.fact["lo_orderdate"][lo3[:m3]])
    # filter_probe_CPU (both filters + probe, compaction) gives the same rows
    lo4, d4 = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    t4 = np.zeros(1, dtype=np.int32)
    lib.oracle_filter_probe_CPU(f, p, Offset(iptr(lo4), None, None, None, iptr(d4)), n, iptr(t4), 0,
                                sptr(sg), CONV_CPU)
    assert int(t4[0]) == m3
    np.testing.assert_array_equal(np.sort(lo4[:m3]), np.sort(lo3[:m3]))
    # filter_probe_aggr_CPU2 over the discount offsets
    res2 = np.zeros(6, dtype=np.int32)
    lib.oracle_filter_probe_aggr_CPU2(Offset(iptr(off1), None, None, None, None), f, p, g, m1, iptr(res2), 0,
                                      CONV_CPU)
    assert oracle.res_scalar(res2) == semantics.scalar_sum(q, data)
    # probe_aggr over the quantity+discount offsets
    res3 = np.zeros(6, dtype=np.int32)
    lib.oracle_probe_aggr_CPU2(Offset(iptr(off2), None, None, None, None), p, g, m2, iptr(res3), 0)
    assert oracle.res_scalar(res3) == semantics.scalar_sum(q, data)


def test_probe_then_group_by_pipeline(star_case):
    """probe_CPU (offsets of every join) then groupByCPU reproduce the fused
    group-by; probe_group_by_CPU2 with dim offsets from an earlier probe too."""
    q, data = star_case
    if q.kind != "groupby":
        pytest.skip("group-by pipeline")
    hts = OQ.build_tables(q, data, CONV_CPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    n = data.n
    sg = OQ.segment_group(n)
    out = [np.zeros(n, dtype=np.int32) for _ in range(5)]
    total = np.zeros(1, dtype=np.int32)
    lib.oracle_probe_CPU(p, Offset(*[iptr(a) for a in out]), n, iptr(total), 0, sptr(sg), CONV_CPU)
    m = int(total[0])
    res = np.zeros(q.total_val * 6, dtype=np.int32)
    lib.oracle_groupByCPU(Offset(*[iptr(a) for a in out]), g, m, iptr(res), CONV_GPU)
    expect = semantics.group_map(q, data)
    # groupByCPU takes group values from every group column, so tables without
    # a group column contribute 0 exactly as in the fused operator
    assert semantics.decode_groups(res) == expect
    # split: probe table 1 first, then the rest in offset mode with dim_off1
    # carried through and its group value read from the dimension column
    p1 = ProbeArgs()
    p1.key_col1, p1.ht1, p1.dim_len1, p1.min_key1 = p.key_col1, p.ht1, p.dim_len1, p.min_key1
    o1 = [np.zeros(n, dtype=np.int32) for _ in range(5)]
    t1 = np.zeros(1, dtype=np.int32)
    lib.oracle_probe_CPU(p1, Offset(*[iptr(a) for a in o1]), n, iptr(t1), 0, sptr(sg), CONV_CPU)
    m1 = int(t1[0])
    prest = ProbeArgs()
    for k in (2, 3, 4):
        for fld in ("key_col", "ht", "dim_len", "min_key"):
            setattr(prest, f"{fld}{k}", getattr(p, f"{fld}{k}"))
    res2 = np.zeros(q.total_val * 6, dtype=np.int32)
    lib.oracle_probe_group_by_CPU2(Offset(iptr(o1[0]), iptr(o1[1]), None, None, None), prest, g, m1,
                                   iptr(res2), 0, CONV_CPU)
    assert semantics.decode_groups(res2) == expect


def test_merge_of_cpu_and_gpu_halves(star_case):
    """merge (CPUProcessing.cu:1786-1819) of a CPU half and a GPU half equals
    the whole query."""
    q, data = star_case
    if q.kind != "groupby":
        pytest.skip("group-by merge")
    hts = OQ.build_tables(q, data, CONV_CPU)
    cpu = OQ.run_fact(q, data, hts, CONV_CPU, segments=[0, 2], n=SEG + (data.n - 2 * SEG))
    gpu = OQ.run_fact(q, data, hts, CONV_GPU, segments=[1], n=SEG)
    lib.oracle_merge(iptr(cpu), iptr(gpu), q.total_val)
    assert semantics.decode_groups(cpu) == semantics.group_map(q, data)


# -------------------------------------------------------- special cases

def test_empty_input_leaves_result_untouched():
    q, data, _ = cases.load_golden("q21_tiny")
    hts = OQ.build_tables(q, data)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    res = np.zeros(q.total_val * 6, dtype=np.int32)
    lib.oracle_probe_group_by_CPU(p, g, 0, iptr(res), 0, sptr(OQ.segment_group(1)), CONV_CPU)
    assert not res.any()


def test_nothing_and_everything_selected():
    q, data, _ = cases.load_golden("q11_tiny")
    hts = OQ.build_tables(q, data)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    sg = sptr(OQ.segment_group(data.n))
    f.compare1, f.compare2 = 100, 200          # discount never in [100, 200]
    res = np.zeros(6, dtype=np.int32)
    lib.oracle_filter_probe_aggr_CPU(f, p, g, data.n, iptr(res), 0, sg, CONV_CPU)
    assert oracle.res_scalar(res) == 0
    # everything: filters wide open and every date in the table
    f.compare1, f.compare2, f.compare3, f.compare4 = -1, 1 << 30, -1, 1 << 30
    d = data.tables["date"]
    ht = np.zeros(2 * q.joins[0].dim_len, dtype=np.int32)
    lib.oracle_build_CPU(FilterArgs(), BuildArgs(iptr(d["d_datekey"]), None, q.joins[0].dim_len,
                                                 q.joins[0].min_key, 0), len(d["d_datekey"]), iptr(ht), 0,
                         sptr(OQ.segment_group(len(d["d_datekey"]))), CONV_CPU)
    p.ht4 = iptr(ht)
    res[:] = 0
    lib.oracle_filter_probe_aggr_CPU(f, p, g, data.n, iptr(res), 0, sg, CONV_CPU)
    a1 = data.fact["lo_extendedprice"].astype(np.int64)
    a2 = data.fact["lo_discount"].astype(np.int64)
    assert oracle.res_scalar(res) == int((a1 * a2).sum())


def test_keys_at_hash_range_ends():
    """Keys min_key and min_key+len-1 land in slots 0 and len-1; key
    min_key+len wraps to slot 0 (HASH = (k-min) % len)."""
    keys = np.array([100, 100 + 9, 100 + 4], dtype=np.int32)
    vals = np.array([7, 8, 9], dtype=np.int32)
    ht = np.zeros(20, dtype=np.int32)
    lib.oracle_build_CPU(FilterArgs(), BuildArgs(iptr(keys), iptr(vals), 10, 100, 0), 3, iptr(ht), 0,
                         sptr(OQ.segment_group(3)), CONV_GPU)
    assert list(ht[0:2]) == [7, 1] and list(ht[18:20]) == [8, 2] and list(ht[8:10]) == [9, 3]
    probe = np.array([100, 109, 110, 104, 105], dtype=np.int32)
    p = ProbeArgs()
    p.key_col4, p.ht4, p.dim_len4, p.min_key4 = iptr(probe), iptr(ht), 10, 100
    lo, d4 = np.zeros(5, dtype=np.int32), np.zeros(5, dtype=np.int32)
    t = np.zeros(1, dtype=np.int32)
    lib.oracle_probe_CPU(p, Offset(iptr(lo), None, None, None, iptr(d4)), 5, iptr(t), 0,
                         sptr(OQ.segment_group(5)), CONV_CPU)
    got = sorted(zip(lo[: t[0]], d4[: t[0]]))
    # 110 wraps onto slot 0 (row 0); 105 hits an empty slot
    assert got == [(0, 0), (1, 1), (2, 0), (3, 2)]


def test_build_minmax():
    d = np.array([5, 3, 9, 7, 1], dtype=np.int32)
    flt = np.array([1, 1, 1, 0, 0], dtype=np.int32)
    ht = np.zeros(20, dtype=np.int32)
    mn = np.array([1 << 30], dtype=np.int32)
    mx = np.array([-(1 << 30)], dtype=np.int32)
    f = _filter_args(flt, (1, 1, 0, 0), (1, 1))
    lib.oracle_build_CPU_minmax(f, BuildArgs(iptr(d), None, 10, 0, 10), 5, iptr(ht), iptr(mn), iptr(mx), 0,
                                sptr(OQ.segment_group(5)), CONV_CPU)
    assert (mn[0], mx[0]) == (3, 9)
    assert ht[2 * 5 + 1] == 1 and ht[2 * 3 + 1] == 2 and ht[2 * 9 + 1] == 3 and ht[2 * 7 + 1] == 0
    rows = np.array([3, 4], dtype=np.int32)
    mn[0], mx[0] = 1 << 30, -(1 << 30)
    lib.oracle_build_CPU_minmax2(iptr(rows), FilterArgs(), BuildArgs(iptr(d), None, 10, 0, 10), 2, iptr(ht),
                                 iptr(mn), iptr(mx), 0, CONV_CPU)
    assert (mn[0], mx[0]) == (1, 7)


# ------------------------------------------- offset-mode and remaining pins
#
# Each check below states the function's meaning with numpy alone: the
# dimension row of a fact key is found by sorting the dimension's keys (no
# hashing), filters are numpy masks, group maps come from the pandas
# statement.  So a wrong hash, probe word, offset position, filter, carried
# dimension row or aggregate in the oracle fails one of them.

def _dim_rows(dim_keys, fact_keys):
    """Row of each fact key in the dimension, or -1 (sort + search, no hash)."""
    order = np.argsort(dim_keys, kind="stable")
    sk = dim_keys[order]
    pos = np.minimum(np.searchsorted(sk, fact_keys), len(sk) - 1)
    return np.where(sk[pos] == fact_keys, order[pos], -1)


def _pairs(*cols):
    a = np.stack([np.asarray(c, dtype=np.int64) for c in cols], axis=1)
    return a[np.lexsort(a.T[::-1])] if len(a) else a


@pytest.mark.parametrize("conv", [CONV_CPU, CONV_GPU])
def test_filter_probe_CPU2_pinned(conv):
    """c2 filter_probe_CPU2 (CPUProcessing.cu:88-168): over a permuted row
    list from start_offset on, the rows passing both BETWEEN filters whose
    orderdate joins the year-filtered date table, with the date row."""
    q, data = cases.star("q11", SEG + 77_777, seed=31)
    hts = OQ.build_tables(q, data, conv)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    rows = np.random.default_rng(5).permutation(data.n)[: data.n // 2].astype(np.int32)
    start = 1000
    n = len(rows) - start
    lo, d4 = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    total = np.zeros(1, dtype=np.int32)
    lib.oracle_filter_probe_CPU2(Offset(iptr(rows), None, None, None, None), f, p,
                                 Offset(iptr(lo), None, None, None, iptr(d4)), n, iptr(total), start, conv)
    m = int(total[0])
    r = rows[start:]
    disc, qty = data.fact["lo_discount"][r], data.fact["lo_quantity"][r]
    date = data.tables["date"]
    drow = _dim_rows(date["d_datekey"], data.fact["lo_orderdate"][r])
    ok = (disc >= 1) & (disc <= 3) & (qty >= 0) & (qty <= 24) & (drow >= 0)
    ok &= date["d_year"][np.maximum(drow, 0)] == 1993
    assert m == int(ok.sum()) > 0
    np.testing.assert_array_equal(_pairs(lo[:m], d4[:m]), _pairs(r[ok], drow[ok]))
    if conv == CONV_GPU:
        # filter_probe_GPU3 (GPUProcessing.h:602-688) without a date table
        # passes the incoming dim_off4 through
        carried = np.arange(len(rows), dtype=np.int32) * 3 + 7
        p.key_col4, p.ht4 = None, None
        total[0] = 0
        lib.oracle_filter_probe_CPU2(Offset(iptr(rows), None, None, None, iptr(carried)), f, p,
                                     Offset(iptr(lo), None, None, None, iptr(d4)), n, iptr(total), start, conv)
        okf = (disc >= 1) & (disc <= 3) & (qty >= 0) & (qty <= 24)
        m = int(total[0])
        assert m == int(okf.sum())
        np.testing.assert_array_equal(_pairs(lo[:m], d4[:m]), _pairs(r[okf], carried[start:][okf]))


def _q4way_filtered():
    import copy
    q, data = cases.star("q4way", SEG + 54_321, seed=41)
    qf = copy.deepcopy(q)
    qf.fact_filters = [("lo_revenue", 1_000_000, 6_000_000), ("lo_supplycost", 60_000, 100_000)]
    return qf, data


@pytest.mark.parametrize("conv", [CONV_CPU, CONV_GPU])
def test_filter_probe_group_by_CPU2_pinned(conv):
    """c8 filter_probe_group_by_CPU2 (CPUProcessing.cu:868-1032): fact
    BETWEEN filters + 4 probes + group-by over a permuted row list from
    start_offset on equals the pandas statement over the same rows; with the
    supplier table not probed, its group value comes from the carried
    dimension row (dim_off1) instead."""
    qf, data = _q4way_filtered()
    hts = OQ.build_tables(qf, data, conv)
    f, p, g, keep = OQ.fact_args(qf, data, hts)
    assert f.filter_col1 and f.filter_col2
    rows = np.random.default_rng(6).permutation(data.n)[: 2 * data.n // 3].astype(np.int32)
    start = 777
    res = np.zeros(qf.total_val * 6, dtype=np.int32)
    lib.oracle_filter_probe_group_by_CPU2(Offset(iptr(rows), None, None, None, None), f, p, g,
                                          len(rows) - start, iptr(res), start, conv)
    expect = semantics.group_map(qf, data, rows[start:])
    assert len(expect) > 100
    assert semantics.decode_groups(res) == expect
    # supplier join done earlier: the list holds only rows that pass it, and
    # dim_off1 carries each row's supplier row
    supp = data.tables["supplier"]
    srow = _dim_rows(supp["s_suppkey"], data.fact["lo_suppkey"][rows])
    sok = (srow >= 0) & (supp["s_region"][np.maximum(srow, 0)] == 1)
    rows2 = np.ascontiguousarray(rows[sok])
    doff1 = np.ascontiguousarray(srow[sok].astype(np.int32))
    p.key_col1, p.ht1 = None, None
    res[:] = 0
    lib.oracle_filter_probe_group_by_CPU2(Offset(iptr(rows2), iptr(doff1), None, None, None), f, p, g,
                                          len(rows2), iptr(res), 0, conv)
    assert semantics.decode_groups(res) == semantics.group_map(qf, data, rows2)


@pytest.mark.parametrize("conv", [CONV_CPU, CONV_GPU])
@pytest.mark.parametrize("with_val", [True, False])
@pytest.mark.parametrize("dim_len,min_key", [(60_000, 1), (50_000, 0)])
def test_build_CPU2_pinned(conv, with_val, dim_len, min_key):
    """c10 build_CPU2 (CPUProcessing.cu:1106-1155): from a permuted dimension
    row list (from start_offset on), slot (key-min) % len holds [value, row+1]
    for each listed row and 0 elsewhere.  With min_key 0 and len = |dim|
    (the engine's dense keys) key |dim| wraps onto slot 0."""
    from mordred_amd import datagen
    part = datagen.make_part(50_000, 9)
    keys, vals = part["p_partkey"], part["p_brand1"]
    rows = np.random.default_rng(7).permutation(50_000)[:20_000].astype(np.int32)
    start = 123
    if min_key == 0:
        rows = rows[rows != 49_999]
        rows[start] = 49_999      # key 50,000 -> slot 0
    ht = np.zeros(2 * dim_len, dtype=np.int32)
    lib.oracle_build_CPU2(iptr(rows), FilterArgs(),
                          BuildArgs(iptr(keys), iptr(vals) if with_val else None, dim_len, min_key, 0),
                          len(rows) - start, iptr(ht), start, conv)
    expect = np.zeros(2 * dim_len, dtype=np.int64)
    r = rows[start:].astype(np.int64)
    h = (keys[r].astype(np.int64) - min_key) % dim_len
    expect[2 * h + 1] = r + 1
    if with_val:
        expect[2 * h] = vals[r]
    np.testing.assert_array_equal(ht, expect)
    if min_key == 0:
        assert ht[1] == 50_000


def test_probe_aggr_CPU_pinned():
    """c15 probe_aggr_CPU (CPUProcessing.cu:1491-1558): sum of int64(a1*a2)
    over the rows of a permuted segment group whose orderdate joins the
    year-filtered date table; no fact filter."""
    q, data = cases.star("q11", 2 * SEG + 123_457, seed=43)
    hts = OQ.build_tables(q, data, CONV_CPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    res = np.zeros(6, dtype=np.int32)
    lib.oracle_probe_aggr_CPU(p, g, data.n, iptr(res), 0, sptr(OQ.segment_group(data.n, [1, 0, 2])))
    date = data.tables["date"]
    drow = _dim_rows(date["d_datekey"], data.fact["lo_orderdate"])
    ok = (drow >= 0) & (date["d_year"][np.maximum(drow, 0)] == 1993)
    a = data.fact["lo_extendedprice"].astype(np.int64) * data.fact["lo_discount"].astype(np.int64)
    assert oracle.res_scalar(res) == int(a[ok].sum()) > 0
    # one segment only: rows 2^20 .. 2^21
    res[:] = 0
    lib.oracle_probe_aggr_CPU(p, g, SEG, iptr(res), 0, sptr(OQ.segment_group(SEG, [1])))
    seg1 = np.zeros(data.n, dtype=bool)
    seg1[SEG:2 * SEG] = True
    assert oracle.res_scalar(res) == int(a[ok & seg1].sum())


def test_row_probe_word_by_convention():
    """Row probes: the GPU kernels test the row word ht[2h+1] only
    (BlockProbeGPU, BlockLibrary.cuh:24-30), the CPU path the whole 64-bit
    slot (CPUProcessing.cu:203-207).  A slot with a value word but no row word
    hits on the CPU (row -1) and misses on the GPU."""
    ht = np.zeros(8, dtype=np.int32)
    ht[0:2] = [5, 0]          # slot 0: value only
    ht[2:4] = [0, 3]          # slot 1: row word only -> row 2
    keys = np.array([0, 1, 2], dtype=np.int32)
    p = ProbeArgs()
    p.key_col4, p.ht4, p.dim_len4, p.min_key4 = iptr(keys), iptr(ht), 4, 0
    got = {}
    for conv in (CONV_CPU, CONV_GPU):
        lo, d4 = np.zeros(3, dtype=np.int32), np.zeros(3, dtype=np.int32)
        t = np.zeros(1, dtype=np.int32)
        lib.oracle_probe_CPU(p, Offset(iptr(lo), None, None, None, iptr(d4)), 3, iptr(t), 0,
                             sptr(OQ.segment_group(3)), conv)
        got[conv] = sorted(zip(lo[: t[0]].tolist(), d4[: t[0]].tolist()))
    assert got[CONV_CPU] == [(0, -1), (1, 2)]
    assert got[CONV_GPU] == [(1, 2)]
