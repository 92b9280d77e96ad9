"""GPU parity: every entry point of include/mordred_hip.h against the oracle.

Integer work, so the bar is bit-exact: result buffers (keys and int64 sums)
and hash tables are compared word for word with the oracle run in the GPU
convention; offset lists are compared as sorted (lo_row, dim rows) tuples
because compaction order depends on atomics (GPUProcessing.h:583-599).
Inputs are seeded star schemas with 2 full segments plus a partial one, loaded
into a segment cache at a shuffled slot order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from oracle import CONV_GPU, BuildArgs, FilterArgs, Offset, ProbeArgs, iptr, lib, sptr
from oracle import queries as OQ
from tests import cases, semantics

pytestmark = pytest.mark.gpu

SEG = oracle.SEGMENT_SIZE
N_STAR = 2 * SEG + 123_457


def _t():
    import torch
    return torch


def dev(a):
    return _t().from_numpy(np.ascontiguousarray(a)).cuda()


def stream():
    return _t().cuda.current_stream().cuda_stream


def sync():
    _t().cuda.synchronize()


@pytest.fixture(scope="module", params=["q11", "q21", "q4way"])
def gstar(request, cuda):
    from mordred_amd import query as Q
    q, data = cases.star(request.param, N_STAR, seed=1234)
    cache = Q.load_cache(data, q, shuffle_seed=7)
    return q, data, cache


def oracle_res(q, data, segments=None, n=None):
    return OQ.run_query(q, data, CONV_GPU, segments=segments, n=n)


# ------------------------------------------------------------ whole queries

@pytest.mark.parametrize("name", cases.golden_names())
def test_golden_query_gpu(cuda, name):
    from mordred_amd import query as Q
    q, data, exp = cases.load_golden(name)
    print(f"[gpu] golden {name}", flush=True)
    cache = Q.load_cache(data, q, shuffle_seed=3)
    gq = Q.GpuQuery(q, data, cache)
    gq.run()
    sync()
    res = gq.result()
    np.testing.assert_array_equal(res, oracle_res(q, data))
    if q.kind == "aggr":
        assert oracle.res_scalar(res) == exp["scalar"]
    else:
        assert semantics.decode_groups(res) == exp["groups"]
    for slot, ht in gq.hts.items():
        np.testing.assert_array_equal(ht.cpu().numpy(), exp["ht"][slot])


def test_star_query_gpu(gstar):
    from mordred_amd import query as Q
    q, data, cache = gstar
    print(f"[gpu] star {q.name} n={data.n}", flush=True)
    gq = Q.GpuQuery(q, data, cache)
    for _ in range(2):          # re-run: reset must clear hash tables and result
        gq.run()
        sync()
        np.testing.assert_array_equal(gq.result(), oracle_res(q, data))
    hts = OQ.build_tables(q, data, CONV_GPU)
    for slot, ht in gq.hts.items():
        np.testing.assert_array_equal(ht.cpu().numpy(), hts[slot])


def test_permuted_segment_group_gpu(gstar):
    """Segment group [1, 0, 2] and a one-segment group [1]."""
    from mordred_amd import query as Q
    q, data, cache = gstar
    gq = Q.GpuQuery(q, data, cache)
    gq.reset(); gq.build()
    gq.probe(n=data.n, sg=Q.seg_group(0, segments=[1, 0, 2]))
    sync()
    np.testing.assert_array_equal(gq.result(), oracle_res(q, data, segments=[1, 0, 2], n=data.n))
    gq.res.zero_()
    gq.probe(n=SEG, sg=Q.seg_group(0, segments=[1]))
    sync()
    np.testing.assert_array_equal(gq.result(), oracle_res(q, data, segments=[1], n=SEG))


def test_group_by_global_atomic_path(gstar):
    """total_val above the LDS budget selects one global atomic per row; a
    larger modulus leaves every index unchanged, so sums must agree."""
    from mordred_amd import query as Q
    q, data, cache = gstar
    if q.kind != "groupby":
        pytest.skip("group-by only")
    import copy
    big = copy.deepcopy(q)
    big.total_val = 20_000
    gq = Q.GpuQuery(big, data, cache)
    gq.run()
    sync()
    res = gq.result()
    ref = oracle_res(q, data)
    np.testing.assert_array_equal(res[: ref.size], ref)
    assert not res[ref.size:].any()


def test_null_segment_group_start_offset(gstar):
    """NULL segment group: segment = (start_offset + i) >> 20."""
    from mordred_amd import query as Q
    q, data, cache = gstar
    gq = Q.GpuQuery(q, data, cache)
    gq.reset(); gq.build()
    from mordred_amd import abi
    fn = "mordred_filter_probe_aggr_GPU2" if q.kind == "aggr" else "mordred_filter_probe_group_by_GPU2"
    abi.call(fn, cache.ptr, abi.ref(gq.fargs), C.pointer(gq.pargs), C.pointer(gq.gargs), SEG,
             gq.res.data_ptr(), SEG, None, stream())
    sync()
    np.testing.assert_array_equal(gq.result(), oracle_res(q, data, segments=[1], n=SEG))


# ------------------------------------------------------- per entry point

def _gpu_query(q, data, cache):
    from mordred_amd import query as Q
    gq = Q.GpuQuery(q, data, cache)
    gq.reset(); gq.build()
    sync()
    return gq


def _alloc(n):
    return _t().zeros(max(1, n), dtype=_t().int32, device="cuda")


def _sorted_rows(cols, m):
    a = np.stack([c[:m] for c in cols], axis=1)
    return a[np.lexsort(a.T[::-1])]


def test_probe_GPU2_and_GPU3(gstar):
    from mordred_amd import abi
    q, data, cache = gstar
    gq = _gpu_query(q, data, cache)
    n = data.n
    outs = [_alloc(n) for _ in range(5)]
    total = _alloc(1)
    o = abi.OffsetGPU(*[t.data_ptr() for t in outs])
    abi.call("mordred_probe_GPU2", cache.ptr, C.pointer(gq.pargs), C.pointer(o), n, total.data_ptr(), 0,
             gq.fact_sg.data_ptr(), stream())
    sync()
    m = int(total.item())
    hts = OQ.build_tables(q, data, CONV_GPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    ro = [np.zeros(n, dtype=np.int32) for _ in range(5)]
    rt = np.zeros(1, dtype=np.int32)
    lib.oracle_probe_CPU(p, Offset(*[iptr(a) for a in ro]), n, iptr(rt), 0, sptr(OQ.segment_group(n)), CONV_GPU)
    assert m == int(rt[0])
    got = _sorted_rows([t.cpu().numpy() for t in outs], m)
    np.testing.assert_array_equal(got, _sorted_rows(ro, m))
    # offset mode: probe only the first joined table from the full row list,
    # carrying the others through from a previous probe
    first = q.joins[0].slot
    p1 = abi.ProbeArgsGPU()
    for fld in ("key_idx", "ht", "dim_len", "min_key"):
        setattr(p1, f"{fld}{first}", getattr(gq.pargs, f"{fld}{first}"))
    in_o = abi.OffsetGPU(*[t.data_ptr() for t in outs])
    outs2 = [_alloc(m) for _ in range(5)]
    total2 = _alloc(1)
    abi.call("mordred_probe_GPU3", cache.ptr, C.pointer(in_o), C.pointer(p1),
             C.pointer(abi.OffsetGPU(*[t.data_ptr() for t in outs2])), m, total2.data_ptr(), stream())
    sync()
    assert int(total2.item()) == m      # every row already passed that probe
    np.testing.assert_array_equal(_sorted_rows([t.cpu().numpy() for t in outs2], m), got)


def test_filter_probe_GPU2_and_GPU3(gstar):
    from mordred_amd import abi
    q, data, cache = gstar
    if q.kind != "aggr":
        pytest.skip("Q1.x compaction")
    gq = _gpu_query(q, data, cache)
    n = data.n
    lo, d4 = _alloc(n), _alloc(n)
    total = _alloc(1)
    o = abi.OffsetGPU(lo.data_ptr(), None, None, None, d4.data_ptr())
    abi.call("mordred_filter_probe_GPU2", cache.ptr, abi.ref(gq.fargs), C.pointer(gq.pargs), C.pointer(o), n,
             total.data_ptr(), 0, gq.fact_sg.data_ptr(), stream())
    sync()
    m = int(total.item())
    hts = OQ.build_tables(q, data, CONV_GPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    rlo, rd4 = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    rt = np.zeros(1, dtype=np.int32)
    lib.oracle_filter_probe_CPU(f, p, Offset(iptr(rlo), None, None, None, iptr(rd4)), n, iptr(rt), 0,
                                sptr(OQ.segment_group(n)), CONV_GPU)
    assert m == int(rt[0]) > 0
    got = _sorted_rows([lo.cpu().numpy(), d4.cpu().numpy()], m)
    np.testing.assert_array_equal(got, _sorted_rows([rlo, rd4], m))
    # GPU3 over every row id: same output
    allrows = dev(np.arange(n, dtype=np.int32))
    lo2, d42 = _alloc(n), _alloc(n)
    total2 = _alloc(1)
    abi.call("mordred_filter_probe_GPU3", cache.ptr, C.pointer(abi.OffsetGPU(allrows.data_ptr(), None, None, None, None)),
             abi.ref(gq.fargs), C.pointer(gq.pargs),
             C.pointer(abi.OffsetGPU(lo2.data_ptr(), None, None, None, d42.data_ptr())), n, total2.data_ptr(),
             stream())
    sync()
    assert int(total2.item()) == m
    np.testing.assert_array_equal(_sorted_rows([lo2.cpu().numpy(), d42.cpu().numpy()], m), got)


@pytest.mark.parametrize("modes", [(0, 0), (1, 1), (2, 2), (1, 2), (2, 0)])
def test_filter_GPU2_and_GPU3(gstar, modes):
    from mordred_amd import abi
    q, data, cache = gstar
    if q.kind != "aggr":
        pytest.skip("uses the Q1.1 fact columns")
    n = data.n
    a, b = data.fact["lo_discount"], data.fact["lo_quantity"]
    cmp = (2, 5, 10, 30)
    fa = abi.FilterArgsGPU()
    fa.filter_idx1, fa.filter_idx2 = cache.idx("lo_discount"), cache.idx("lo_quantity")
    fa.compare1, fa.compare2, fa.compare3, fa.compare4 = cmp
    fa.mode1, fa.mode2 = modes
    out, total = _alloc(n), _alloc(1)
    sg = _t().tensor(list(range((n + SEG - 1) // SEG)), dtype=_t().int16, device="cuda")
    abi.call("mordred_filter_GPU2", cache.ptr, C.pointer(fa), out.data_ptr(), n, total.data_ptr(), 0,
             sg.data_ptr(), stream())
    sync()
    m = int(total.item())
    rf = FilterArgs()
    rf.filter_col1, rf.filter_col2 = iptr(a), iptr(b)
    rf.compare1, rf.compare2, rf.compare3, rf.compare4 = cmp
    rf.mode1, rf.mode2 = modes
    rout, rt = np.zeros(n, dtype=np.int32), np.zeros(1, dtype=np.int32)
    lib.oracle_filter_CPU(rf, iptr(rout), n, iptr(rt), 0, sptr(OQ.segment_group(n)), CONV_GPU)
    assert m == int(rt[0])
    got = np.sort(out.cpu().numpy()[:m])
    np.testing.assert_array_equal(got, np.sort(rout[:m]))
    # GPU3 from a row list that skips every third row
    rows = np.arange(0, n, 3, dtype=np.int32)
    drows = dev(rows)
    out3, t3 = _alloc(len(rows)), _alloc(1)
    abi.call("mordred_filter_GPU3", cache.ptr, drows.data_ptr(), C.pointer(fa), out3.data_ptr(), len(rows),
             t3.data_ptr(), stream())
    sync()
    rout3, rt3 = np.zeros(len(rows), dtype=np.int32), np.zeros(1, dtype=np.int32)
    lib.oracle_filter_CPU2(iptr(rows), rf, iptr(rout3), len(rows), iptr(rt3), 0, CONV_GPU)
    m3 = int(t3.item())
    assert m3 == int(rt3[0])
    np.testing.assert_array_equal(np.sort(out3.cpu().numpy()[:m3]), np.sort(rout3[:m3]))


def test_offset_mode_group_by_family(gstar):
    """probe_group_by_GPU3, filter_probe_group_by_GPU3 and groupByGPU from the
    offset lists of probe_GPU2 reproduce the fused group-by."""
    from mordred_amd import abi
    q, data, cache = gstar
    if q.kind != "groupby":
        pytest.skip("group-by only")
    gq = _gpu_query(q, data, cache)
    n = data.n
    ref = oracle_res(q, data)
    # (a) groupByGPU over the offsets of all probes
    outs = [_alloc(n) for _ in range(5)]
    total = _alloc(1)
    o = abi.OffsetGPU(*[t.data_ptr() for t in outs])
    abi.call("mordred_probe_GPU2", cache.ptr, C.pointer(gq.pargs), C.pointer(o), n, total.data_ptr(), 0,
             gq.fact_sg.data_ptr(), stream())
    sync()
    m = int(total.item())
    res = _alloc(q.total_val * 6)
    abi.call("mordred_groupByGPU", cache.ptr, C.pointer(o), C.pointer(gq.gargs), m, res.data_ptr(), stream())
    sync()
    np.testing.assert_array_equal(res.cpu().numpy(), ref)
    # (b) probe the first table only, then probe_group_by_GPU3 on the rest with
    # the first table's group value read through dim_off
    first = q.joins[0].slot
    p1 = abi.ProbeArgsGPU()
    prest = abi.ProbeArgsGPU()
    for fld in ("key_idx", "ht", "dim_len", "min_key"):
        setattr(p1, f"{fld}{first}", getattr(gq.pargs, f"{fld}{first}"))
        for j in q.joins[1:]:
            setattr(prest, f"{fld}{j.slot}", getattr(gq.pargs, f"{fld}{j.slot}"))
    outs1 = [_alloc(n) for _ in range(5)]
    t1 = _alloc(1)
    o1 = abi.OffsetGPU(*[t.data_ptr() for t in outs1])
    abi.call("mordred_probe_GPU2", cache.ptr, C.pointer(p1), C.pointer(o1), n, t1.data_ptr(), 0,
             gq.fact_sg.data_ptr(), stream())
    sync()
    m1 = int(t1.item())
    for fn in ("mordred_probe_group_by_GPU3", "mordred_filter_probe_group_by_GPU3"):
        res.zero_()
        if fn.startswith("mordred_filter"):
            abi.call(fn, cache.ptr, C.pointer(o1), None, C.pointer(prest), C.pointer(gq.gargs), m1,
                     res.data_ptr(), stream())
        else:
            abi.call(fn, cache.ptr, C.pointer(o1), C.pointer(prest), C.pointer(gq.gargs), m1, res.data_ptr(),
                     stream())
        sync()
        np.testing.assert_array_equal(res.cpu().numpy(), ref, err_msg=fn)


def test_offset_mode_aggregate_family(gstar):
    """aggregationGPU, probe_aggr_GPU3 and filter_probe_aggr_GPU3 (which
    re-checks filter 2 only) against the oracle."""
    from mordred_amd import abi
    q, data, cache = gstar
    if q.kind != "aggr":
        pytest.skip("Q1.x only")
    gq = _gpu_query(q, data, cache)
    n = data.n
    expect = semantics.scalar_sum(q, data)
    hts = OQ.build_tables(q, data, CONV_GPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    # rows passing filter 1 (discount) -> offset list
    d = data.fact["lo_discount"]
    rows1 = np.nonzero((d >= 1) & (d <= 3))[0].astype(np.int32)
    r1 = dev(rows1)
    o = abi.OffsetGPU(r1.data_ptr(), None, None, None, None)
    res = _alloc(6)
    abi.call("mordred_filter_probe_aggr_GPU3", cache.ptr, C.pointer(o), abi.ref(gq.fargs), C.pointer(gq.pargs),
             C.pointer(gq.gargs), len(rows1), res.data_ptr(), stream())
    sync()
    assert oracle.res_scalar(res.cpu().numpy()) == expect
    # every row id: GPU3 re-checks filter 2 only, so filter 1 is not applied
    allr = np.arange(n, dtype=np.int32)
    rall = dev(allr)
    res.zero_()
    abi.call("mordred_filter_probe_aggr_GPU3", cache.ptr, C.pointer(abi.OffsetGPU(rall.data_ptr(), None, None, None, None)),
             abi.ref(gq.fargs), C.pointer(gq.pargs), C.pointer(gq.gargs), n, res.data_ptr(), stream())
    sync()
    rres = np.zeros(6, dtype=np.int32)
    lib.oracle_filter_probe_aggr_CPU2(Offset(iptr(allr), None, None, None, None), f, p, g, n, iptr(rres), 0, CONV_GPU)
    assert oracle.res_scalar(res.cpu().numpy()) == oracle.res_scalar(rres)
    # probe_aggr_GPU3 over rows passing both filters
    qn = data.fact["lo_quantity"]
    rows2 = np.nonzero((d >= 1) & (d <= 3) & (qn >= 0) & (qn <= 24))[0].astype(np.int32)
    r2 = dev(rows2)
    res.zero_()
    abi.call("mordred_probe_aggr_GPU3", cache.ptr, C.pointer(abi.OffsetGPU(r2.data_ptr(), None, None, None, None)),
             C.pointer(gq.pargs), C.pointer(gq.gargs), len(rows2), res.data_ptr(), stream())
    sync()
    assert oracle.res_scalar(res.cpu().numpy()) == expect
    # probe_aggr_GPU2: probe only (no filters) over the segment group
    res.zero_()
    abi.call("mordred_probe_aggr_GPU2", cache.ptr, C.pointer(gq.pargs), C.pointer(gq.gargs), n, res.data_ptr(), 0,
             gq.fact_sg.data_ptr(), stream())
    sync()
    rres[:] = 0
    lib.oracle_probe_aggr_CPU(p, g, n, iptr(rres), 0, sptr(OQ.segment_group(n)))
    assert oracle.res_scalar(res.cpu().numpy()) == oracle.res_scalar(rres)
    # aggregationGPU over the rows of a probe: filter_probe_GPU2 then aggregate
    lo, total = _alloc(n), _alloc(1)
    abi.call("mordred_filter_probe_GPU2", cache.ptr, abi.ref(gq.fargs), C.pointer(gq.pargs),
             C.pointer(abi.OffsetGPU(lo.data_ptr(), None, None, None, None)), n, total.data_ptr(), 0,
             gq.fact_sg.data_ptr(), stream())
    sync()
    res.zero_()
    abi.call("mordred_aggregationGPU", cache.ptr, lo.data_ptr(), C.pointer(gq.gargs), int(total.item()),
             res.data_ptr(), stream())
    sync()
    assert oracle.res_scalar(res.cpu().numpy()) == expect


def test_build_GPU3_and_build_GPU(gstar):
    """Offset-list build and on-demand build produce the oracle's table."""
    from mordred_amd import abi
    q, data, cache = gstar
    j = q.joins[-1] if q.kind == "aggr" else q.joins[1]
    t = data.tables[j.table]
    m = len(t[j.dim_key])
    # rows passing the dimension filter, as the filter kernel would produce
    if j.filter_col:
        x = t[j.filter_col]
        keep = (x >= j.c1) & (x <= j.c2)
    else:
        keep = np.ones(m, dtype=bool)
    rows = np.nonzero(keep)[0].astype(np.int32)
    ref = np.zeros(2 * j.dim_len, dtype=np.int32)
    b = BuildArgs(iptr(t[j.dim_key]), iptr(t[j.group_col]) if j.group_col else None, j.dim_len, j.min_key, 0)
    lib.oracle_build_CPU2(iptr(rows), FilterArgs(), b, len(rows), iptr(ref), 0, CONV_GPU)
    ht = _alloc(2 * j.dim_len)
    drows = dev(rows)
    bg = abi.BuildArgsGPU(cache.idx(j.dim_key), cache.idx(j.group_col), j.dim_len, j.min_key, 0)
    abi.call("mordred_build_GPU3", cache.ptr, drows.data_ptr(), None, C.pointer(bg), len(rows), ht.data_ptr(),
             stream())
    sync()
    np.testing.assert_array_equal(ht.cpu().numpy(), ref)
    # on-demand build of one segment from plain columns (segment_index 0)
    ht.zero_()
    keyc, valc = dev(t[j.dim_key]), (dev(t[j.group_col]) if j.group_col else None)
    fc = dev(t[j.filter_col]) if j.filter_col else None
    abi.call("mordred_build_GPU", keyc.data_ptr(), valc.data_ptr() if valc is not None else None,
             fc.data_ptr() if fc is not None else None, j.c1, j.c2, 1, min(m, SEG), ht.data_ptr(), j.dim_len,
             j.min_key, 0, stream())
    sync()
    ref2 = np.zeros(2 * j.dim_len, dtype=np.int32)
    fa = FilterArgs()
    if j.filter_col:
        fa.filter_col1, fa.compare1, fa.compare2, fa.mode1 = iptr(t[j.filter_col]), j.c1, j.c2, 1
    lib.oracle_build_CPU(fa, b, min(m, SEG), iptr(ref2), 0, sptr(OQ.segment_group(min(m, SEG))), CONV_GPU)
    np.testing.assert_array_equal(ht.cpu().numpy(), ref2)


def test_on_demand_fact_kernels(gstar):
    """filter_probe_aggr_GPU and probe_group_by_GPU over one plain segment."""
    from mordred_amd import abi
    q, data, cache = gstar
    gq = _gpu_query(q, data, cache)
    s0 = 1   # second segment of the fact table, as plain column pointers
    cols = {c: dev(data.fact[c][s0 * SEG:(s0 + 1) * SEG]) for c in q.fact_columns()}
    ptr = lambda c: cols[c].data_ptr() if c else None  # noqa: E731
    res = _alloc(max(1, q.total_val) * 6)
    if q.kind == "aggr":
        (c1, l1, h1), (c2, l2, h2) = q.fact_filters
        j = q.joins[0]
        abi.call("mordred_filter_probe_aggr_GPU", ptr(c1), ptr(c2), l1, h1, l2, h2, 1, 1, ptr(j.fact_key),
                 gq.hts[j.slot].data_ptr(), j.dim_len, j.min_key, ptr(q.aggr[0]), ptr(q.aggr[1]), SEG,
                 res.data_ptr(), stream())
    else:
        args = []
        for k in range(1, 5):
            jj = q.join(k)
            args.append(ptr(jj.fact_key) if jj else None)
        for k in range(1, 5):
            jj = q.join(k)
            args += [gq.hts[k].data_ptr() if jj else None, jj.dim_len if jj else 0]
        for k in range(1, 5):
            jj = q.join(k)
            args.append(jj.min_key if jj else 0)
        args += [ptr(q.aggr[0]), ptr(q.aggr[1])]
        for k in range(1, 5):
            jj = q.join(k)
            args.append(jj.min_val if jj and jj.group_col else 0)
        for k in range(1, 5):
            jj = q.join(k)
            args.append(jj.unique_val if jj and jj.group_col else 0)
        args += [q.total_val, SEG, res.data_ptr(), stream()]
        abi.call("mordred_probe_group_by_GPU", *args)
    sync()
    np.testing.assert_array_equal(res.cpu().numpy(), oracle_res(q, data, segments=[s0], n=SEG))


@pytest.fixture(scope="module")
def big_part(cuda):
    """Part table above the compaction threshold (300,000 slots) with enough
    fact rows (>= 16 per slot) that the fused kernel probes compact tables:
    supplier 1 bit in LDS, part 16-bit brand entries in L2; row probes of the
    part table need 19 bits and fall back to the hash table itself."""
    from mordred_amd import query as Q
    q = Q.q21(5_000, 300_000)
    data = Q.make_data(q, 5 * SEG + 4321, seed=77, S=5_000, P=300_000)
    cache = Q.load_cache(data, q, shuffle_seed=11)
    return q, data, cache


def test_compact_tables_query(big_part):
    from mordred_amd import query as Q
    q, data, cache = big_part
    print("[gpu] compact tables: fused group-by", flush=True)
    gq = Q.GpuQuery(q, data, cache)
    for _ in range(2):
        gq.run()
        sync()
        np.testing.assert_array_equal(gq.result(), oracle_res(q, data))


def test_compact_tables_probe_rows(big_part):
    from mordred_amd import abi
    q, data, cache = big_part
    print("[gpu] compact tables: probe_GPU2 rows", flush=True)
    gq = _gpu_query(q, data, cache)
    n = data.n
    outs = [_alloc(n) for _ in range(5)]
    total = _alloc(1)
    o = abi.OffsetGPU(*[t.data_ptr() for t in outs])
    abi.call("mordred_probe_GPU2", cache.ptr, C.pointer(gq.pargs), C.pointer(o), n, total.data_ptr(), 0,
             gq.fact_sg.data_ptr(), stream())
    sync()
    m = int(total.item())
    hts = OQ.build_tables(q, data, CONV_GPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    ro = [np.zeros(n, dtype=np.int32) for _ in range(5)]
    rt = np.zeros(1, dtype=np.int32)
    lib.oracle_probe_CPU(p, Offset(*[iptr(a) for a in ro]), n, iptr(rt), 0, sptr(OQ.segment_group(n)), CONV_GPU)
    assert m == int(rt[0]) > 0
    np.testing.assert_array_equal(_sorted_rows([t.cpu().numpy() for t in outs], m), _sorted_rows(ro, m))


@pytest.mark.parametrize("kind", ["q21", "q4way"])
def test_one_fact_filter_group_by(cuda, kind):
    """The fused group-by with only the first fact filter (a revenue range):
    its own specialised shape, over a long enough scan that the dimension
    tables are compacted and the survivors go through the ring."""
    import dataclasses
    from mordred_amd import query as Q
    base, data = cases.star(kind, 5 * SEG + 777, seed=55)
    q = dataclasses.replace(base, name=base.name + "_f1", fact_filters=[("lo_revenue", 1_000_000, 6_000_000)])
    cache = Q.load_cache(data, q, shuffle_seed=17)
    print(f"[gpu] one fact filter: {q.name}", flush=True)
    gq = Q.GpuQuery(q, data, cache)
    gq.run()
    sync()
    res = gq.result()
    np.testing.assert_array_equal(res, oracle_res(q, data))
    assert semantics.decode_groups(res) == semantics.group_map(q, data)


def test_lds_staged_tables_and_late_columns(cuda):
    """A fact scan long enough (>= 256 rows per date slot) that the date table
    is compacted and staged in LDS next to the supplier table, with the part
    table's existence bitmap tested per tile and the date keys and revenue
    gathered by row for the survivors; grouped by (brand, year)."""
    from mordred_amd import query as Q
    q = Q.q21(5_000, 300_000)
    n = 16 * SEG + 9_999
    assert n >= 256 * q.join(4).dim_len
    data = Q.make_data(q, n, seed=78, S=5_000, P=300_000)
    cache = Q.load_cache(data, q, shuffle_seed=12)
    print("[gpu] LDS-staged tables: fused group-by", flush=True)
    gq = Q.GpuQuery(q, data, cache)
    ref = oracle_res(q, data)
    for _ in range(2):
        gq.run()
        sync()
        np.testing.assert_array_equal(gq.result(), ref)


@pytest.mark.parametrize("case", ["brands_past_range", "strides_not_mixed_radix"])
def test_group_keys_decoded_or_stored(cuda, case):
    """Group keys in the LDS group-by sink.  When the strides form a mixed
    radix (Q2.1: year stride 1, brand stride 7, 7,000 groups), rows whose
    digits lie in range leave their keys to the flush, which decodes them
    from the group index; here the brands are moved to 500..1499, so rows of
    brands 1000-1499 wrap past the index's range and store their keys
    themselves.  No two key tuples share an index (brands 0-499 are absent),
    so the result is deterministic.  With 7,001 groups the strides are not a
    mixed radix and every first touch stores its keys."""
    import copy
    from mordred_amd import query as Q
    q, data = cases.star("q21", N_STAR, seed=77)
    print(f"[gpu] group keys {case}", flush=True)
    data.tables["part"] = dict(data.tables["part"])
    if case == "brands_past_range":
        data.tables["part"]["p_brand1"] = (data.tables["part"]["p_brand1"] % 1000 + 500).astype(np.int32)
    else:
        q = copy.deepcopy(q)
        q.total_val = 7001
    cache = Q.load_cache(data, q, shuffle_seed=5)
    gq = Q.GpuQuery(q, data, cache)
    gq.run()
    sync()
    res = gq.result()
    ref = oracle_res(q, data)
    np.testing.assert_array_equal(res, ref)
    groups = ref.reshape(-1, 6)
    touched = groups[:, 4:].any(axis=1)
    if case == "brands_past_range":
        # both kinds of group are present
        assert (groups[touched, 2] >= 1000).any() and (groups[touched, 2] < 1000).any()


# ------------------------------------- offset-mode kernels with fact filters

def _fact_filter_args(cache, spec):
    """filterArgsGPU of up to two fact BETWEEN filters [(col or None, lo, hi)]."""
    from mordred_amd import abi
    fa = abi.FilterArgsGPU()
    (c1, l1, h1), (c2, l2, h2) = spec
    fa.filter_idx1 = cache.idx(c1) if c1 else None
    fa.filter_idx2 = cache.idx(c2) if c2 else None
    fa.compare1, fa.compare2, fa.compare3, fa.compare4 = l1, h1, l2, h2
    fa.mode1 = fa.mode2 = 1
    return fa


@pytest.mark.parametrize("spec", ["both", "first", "second"])
def test_filter_probe_group_by_GPU3_with_fact_filters(gstar, spec):
    """filter_probe_group_by_GPU3 (GPUProcessing.h:157-272) with fact filters
    (:187-197): the offset list of a first probe_GPU2 (table 1 only), then
    the BETWEEN filters, the remaining probes and the group-by, word for word
    against oracle_filter_probe_group_by_CPU2 in the GPU convention fed the
    same lists.  One filter column, the other, and both."""
    from mordred_amd import abi
    q, data, cache = gstar
    if q.kind != "groupby":
        pytest.skip("group-by only")
    print(f"[gpu] k10 filters {q.name} {spec}", flush=True)
    gq = _gpu_query(q, data, cache)
    n = data.n
    f1 = ("lo_revenue", 1_000_000, 6_000_000)
    f2 = ("lo_orderdate", 19930000, 19951231)
    none = (None, 0, 0)
    fs = {"both": (f1, f2), "first": (f1, none), "second": (none, f2)}[spec]
    fa = _fact_filter_args(cache, fs)
    first = q.joins[0].slot
    p1, prest = abi.ProbeArgsGPU(), abi.ProbeArgsGPU()
    for fld in ("key_idx", "ht", "dim_len", "min_key"):
        setattr(p1, f"{fld}{first}", getattr(gq.pargs, f"{fld}{first}"))
        for j in q.joins[1:]:
            setattr(prest, f"{fld}{j.slot}", getattr(gq.pargs, f"{fld}{j.slot}"))
    outs1 = [_alloc(n) for _ in range(5)]
    t1 = _alloc(1)
    o1 = abi.OffsetGPU(*[t.data_ptr() for t in outs1])
    abi.call("mordred_probe_GPU2", cache.ptr, C.pointer(p1), C.pointer(o1), n, t1.data_ptr(), 0,
             gq.fact_sg.data_ptr(), stream())
    sync()
    m1 = int(t1.item())
    res = _alloc(q.total_val * 6)
    abi.call("mordred_filter_probe_group_by_GPU3", cache.ptr, C.pointer(o1), C.pointer(fa), C.pointer(prest),
             C.pointer(gq.gargs), m1, res.data_ptr(), stream())
    sync()
    # the oracle over the same lists
    hts = OQ.build_tables(q, data, CONV_GPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    for k in (first,):
        setattr(p, f"key_col{k}", None)
        setattr(p, f"ht{k}", None)
    rf = FilterArgs()
    rf.filter_col1 = iptr(data.fact[fs[0][0]]) if fs[0][0] else None
    rf.filter_col2 = iptr(data.fact[fs[1][0]]) if fs[1][0] else None
    rf.compare1, rf.compare2, rf.compare3, rf.compare4 = fs[0][1], fs[0][2], fs[1][1], fs[1][2]
    rf.mode1 = rf.mode2 = 1
    lo = outs1[0].cpu().numpy()[:m1].copy()
    dims = [t.cpu().numpy()[:m1].copy() for t in outs1[1:]]
    ref = np.zeros(q.total_val * 6, dtype=np.int32)
    lib.oracle_filter_probe_group_by_CPU2(Offset(iptr(lo), *[iptr(d) for d in dims]), rf, p, g, m1, iptr(ref), 0,
                                          CONV_GPU)
    got = res.cpu().numpy()
    assert oracle.res_sums(ref).any()
    np.testing.assert_array_equal(got, ref)
    # and the filters really cut rows: the unfiltered offset-mode call differs
    if spec == "both":
        assert not np.array_equal(got, oracle_res(q, data))


def test_filter_probe_GPU3_permuted_list(gstar):
    """filter_probe_GPU3 (GPUProcessing.h:602-688) over a permuted half of the
    row ids against oracle_filter_probe_CPU2 (GPU convention); without a date
    table it passes the incoming dim_off4 through."""
    from mordred_amd import abi
    q, data, cache = gstar
    if q.kind != "aggr":
        pytest.skip("Q1.x compaction")
    gq = _gpu_query(q, data, cache)
    rows = np.random.default_rng(17).permutation(data.n)[: data.n // 2].astype(np.int32)
    carried = (np.arange(len(rows), dtype=np.int32) * 5 + 3).astype(np.int32)
    m_in = len(rows)
    drows, dcar = dev(rows), dev(carried)
    hts = OQ.build_tables(q, data, CONV_GPU)
    f, p, g, keep = OQ.fact_args(q, data, hts)
    for with_table in (True, False):
        pa = gq.pargs
        if not with_table:
            pa = abi.ProbeArgsGPU()
            p.key_col4, p.ht4 = None, None
        lo, d4, total = _alloc(m_in), _alloc(m_in), _alloc(1)
        abi.call("mordred_filter_probe_GPU3", cache.ptr,
                 C.pointer(abi.OffsetGPU(drows.data_ptr(), None, None, None, dcar.data_ptr())),
                 abi.ref(gq.fargs), C.pointer(pa), C.pointer(abi.OffsetGPU(lo.data_ptr(), None, None, None,
                                                                           d4.data_ptr())),
                 m_in, total.data_ptr(), stream())
        sync()
        m = int(total.item())
        rlo, rd4, rt = np.zeros(m_in, dtype=np.int32), np.zeros(m_in, dtype=np.int32), np.zeros(1, dtype=np.int32)
        lib.oracle_filter_probe_CPU2(Offset(iptr(rows), None, None, None, iptr(carried)), f, p,
                                     Offset(iptr(rlo), None, None, None, iptr(rd4)), m_in, iptr(rt), 0, CONV_GPU)
        assert m == int(rt[0]) > 0, with_table
        np.testing.assert_array_equal(_sorted_rows([lo.cpu().numpy(), d4.cpu().numpy()], m),
                                      _sorted_rows([rlo, rd4], m))


@pytest.mark.parametrize("val_max", [0, 1 << 30])
def test_build_GPU_minmax_and_minmax2(gstar, val_max):
    """build_GPU_minmax (GPUProcessing.h:1030-1113) and build_GPU_minmax2
    (:1115-1172): the hash table word for word and the merged key min / max
    against oracle_build_CPU_minmax / _minmax2 in the GPU convention.  With
    val_max = 0 (what the reference's callers put there,
    CPUGPUProcessing.cu:1606-1609) every workgroup's starting minimum 0 wins;
    with a large val_max the minimum is the smallest built key."""
    from mordred_amd import abi
    q, data, cache = gstar
    j = q.joins[-1] if q.kind == "aggr" else q.joins[1]
    t = data.tables[j.table]
    m = len(t[j.dim_key])
    print(f"[gpu] build minmax {q.name} {j.table} val_max={val_max}", flush=True)
    bg = abi.BuildArgsGPU(cache.idx(j.dim_key), cache.idx(j.group_col), j.dim_len, j.min_key, val_max)
    fa = None
    rf = FilterArgs()
    if j.filter_col:
        fa = abi.FilterArgsGPU()
        fa.filter_idx1, fa.compare1, fa.compare2, fa.mode1 = cache.idx(j.filter_col), j.c1, j.c2, j.filter_mode
        rf.filter_col1, rf.compare1, rf.compare2, rf.mode1 = iptr(t[j.filter_col]), j.c1, j.c2, j.filter_mode
    b = BuildArgs(iptr(t[j.dim_key]), iptr(t[j.group_col]) if j.group_col else None, j.dim_len, j.min_key, val_max)
    big, small = np.int32(2**31 - 1), np.int32(-2**31)
    ht = _alloc(2 * j.dim_len)
    mm = dev(np.array([big, small], dtype=np.int32))
    sg = dev(OQ.segment_group(m))
    abi.call("mordred_build_GPU_minmax", cache.ptr, abi.ref(fa), C.pointer(bg), m, ht.data_ptr(), mm.data_ptr(),
             mm.data_ptr() + 4, 0, sg.data_ptr(), stream())
    sync()
    ref = np.zeros(2 * j.dim_len, dtype=np.int32)
    rmn, rmx = np.array([big], dtype=np.int32), np.array([small], dtype=np.int32)
    lib.oracle_build_CPU_minmax(rf, b, m, iptr(ref), iptr(rmn), iptr(rmx), 0, sptr(OQ.segment_group(m)), CONV_GPU)
    np.testing.assert_array_equal(ht.cpu().numpy(), ref)
    got = mm.cpu().numpy()
    assert (int(got[0]), int(got[1])) == (int(rmn[0]), int(rmx[0]))
    if val_max:
        keys = t[j.dim_key]
        if j.filter_col:
            x = t[j.filter_col]
            keys = keys[(x >= j.c1) & (x <= j.c2)]
        assert (int(got[0]), int(got[1])) == (int(keys.min()), int(keys.max()))
    # offset-list twin over a permuted half of the dimension rows, no filter
    rows = np.random.default_rng(23).permutation(m)[: m // 2].astype(np.int32)
    ht.zero_()
    mm.copy_(dev(np.array([big, small], dtype=np.int32)))
    drows = dev(rows)
    abi.call("mordred_build_GPU_minmax2", cache.ptr, drows.data_ptr(), None, C.pointer(bg), len(rows), ht.data_ptr(),
             mm.data_ptr(), mm.data_ptr() + 4, stream())
    sync()
    ref[:] = 0
    rmn[0], rmx[0] = big, small
    lib.oracle_build_CPU_minmax2(iptr(rows), FilterArgs(), b, len(rows), iptr(ref), iptr(rmn), iptr(rmx), 0, CONV_GPU)
    np.testing.assert_array_equal(ht.cpu().numpy(), ref)
    got = mm.cpu().numpy()
    assert (int(got[0]), int(got[1])) == (int(rmn[0]), int(rmx[0]))
    if val_max:
        keys = t[j.dim_key][rows]
        assert (int(got[0]), int(got[1])) == (int(keys.min()), int(keys.max()))
