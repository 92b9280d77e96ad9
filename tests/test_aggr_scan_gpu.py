"""GPU parity of the segment-mode aggregate scan: filter_probe_aggr_GPU2 and
probe_aggr_GPU2 (GPUProcessing.h:1693-1775) in every specialised shape.

Shapes: no filter, filter 1 only, filter 2 only, both filters, each with and
without the date probe; aggregate pairs whose second column is filter column
1 (the column is then read once), given in either order, or no filter column
at all.  Also: the last, partial tile at several lengths, a permuted segment
group, a cache that is not 16-byte aligned (element loads instead of 128-bit
ones), and the compact date table staged in LDS (a probe-only scan, and a
filtered scan over 16M rows).

The expected sum is the closed form of the query over the fact columns
(d_year = datekey / 10000, so the date join is a test on lo_orderdate),
computed with numpy; integer work, so every sum must be equal.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from tests import cases

pytestmark = pytest.mark.gpu

SEG = oracle.SEGMENT_SIZE
N = 2 * SEG + 5_000      # 5,000 rows past two segments: a partial last tile
YEAR = 1993

FILTERS = {
    "none": (None, None),
    "f1": (("lo_discount", 1, 3), None),
    "f2": (None, ("lo_quantity", 0, 24)),
    "both": (("lo_discount", 1, 3), ("lo_quantity", 0, 24)),
}
AGGRS = [("lo_extendedprice", "lo_discount"), ("lo_discount", "lo_extendedprice"),
         ("lo_quantity", "lo_extendedprice")]


def _t():
    import torch
    return torch


def sync():
    _t().cuda.synchronize()


def stream():
    return _t().cuda.current_stream().cuda_stream


@pytest.fixture(scope="module")
def star(cuda):
    from mordred_amd import query as Q
    q, data = cases.star("q11", N, seed=99)
    cache = Q.load_cache(data, q, shuffle_seed=5)
    gq = Q.GpuQuery(q, data, cache)
    gq.reset()
    gq.build()
    sync()
    return q, data, cache, gq


def expected(data, filters, probe, aggr, rows=None):
    f = data.fact
    idx = np.arange(data.n) if rows is None else np.asarray(rows)
    keep = np.ones(len(idx), dtype=bool)
    for flt in filters:
        if flt:
            c, lo, hi = flt
            keep &= (f[c][idx] >= lo) & (f[c][idx] <= hi)
    if probe:
        keep &= f["lo_orderdate"][idx] // 10000 == YEAR
    sel = idx[keep]
    # every product fits int32 here (<= 50 * 10.5M), as in the int32 kernel
    return int((f[aggr[0]][sel].astype(np.int64) * f[aggr[1]][sel].astype(np.int64)).sum())


def run_aggr(cache_ptr, cache, gq, filters, probe, aggr, n, sg):
    from mordred_amd import abi
    f = None
    if filters[0] or filters[1]:
        f = abi.FilterArgsGPU()
        if filters[0]:
            c, lo, hi = filters[0]
            f.filter_idx1, f.compare1, f.compare2, f.mode1 = cache.idx(c), lo, hi, 1
        if filters[1]:
            c, lo, hi = filters[1]
            f.filter_idx2, f.compare3, f.compare4, f.mode2 = cache.idx(c), lo, hi, 1
    g = abi.GroupbyArgsGPU()
    g.aggr_idx1, g.aggr_idx2, g.total_val = cache.idx(aggr[0]), cache.idx(aggr[1]), 1
    res = _t().zeros(6, dtype=_t().int32, device="cuda")
    abi.call("mordred_filter_probe_aggr_GPU2", cache_ptr, abi.ref(f), C.pointer(gq.pargs) if probe else None,
             C.pointer(g), n, res.data_ptr(), 0, sg.data_ptr(), stream())
    sync()
    return oracle.res_scalar(res.cpu().numpy())


@pytest.mark.parametrize("aggr", AGGRS, ids=["price*disc", "disc*price", "qty*price"])
@pytest.mark.parametrize("probe", [False, True], ids=["noprobe", "probe"])
@pytest.mark.parametrize("fname", list(FILTERS))
def test_aggr_scan_shapes(star, fname, probe, aggr):
    q, data, cache, gq = star
    print(f"[gpu] aggr scan {fname} probe={probe} {aggr}", flush=True)
    got = run_aggr(cache.ptr, cache, gq, FILTERS[fname], probe, aggr, data.n, gq.fact_sg)
    assert got == expected(data, FILTERS[fname], probe, aggr)


@pytest.mark.parametrize("n", [1, 4_095, 4_097, SEG + 1, 2 * SEG])
def test_aggr_scan_partial_tiles(star, n):
    from mordred_amd import query as Q
    q, data, cache, gq = star
    print(f"[gpu] aggr scan n={n}", flush=True)
    sg = Q.seg_group(n)
    flt = FILTERS["both"]
    aggr = AGGRS[0]
    assert run_aggr(cache.ptr, cache, gq, flt, True, aggr, n, sg) == expected(data, flt, True, aggr, np.arange(n))


def test_aggr_scan_permuted_segment_group(star):
    """Segment group [1, 0, 2]: the two full segments in swapped order, then
    the short one, so every row is scanned once."""
    from mordred_amd import query as Q
    q, data, cache, gq = star
    print("[gpu] aggr scan segment group [1, 0, 2]", flush=True)
    sg = Q.seg_group(0, segments=[1, 0, 2])
    flt = FILTERS["both"]
    assert run_aggr(cache.ptr, cache, gq, flt, True, AGGRS[0], data.n, sg) == expected(data, flt, True, AGGRS[0])


def test_aggr_scan_unaligned_cache(star):
    """A cache base that is 4 bytes past a 16-byte boundary: the scan reads
    the columns element by element (same answers)."""
    q, data, cache, gq = star
    print("[gpu] aggr scan, unaligned cache", flush=True)
    torch = _t()
    buf = torch.zeros(cache.buf.numel() + 4, dtype=torch.int32, device="cuda")
    buf[1:1 + cache.buf.numel()].copy_(cache.buf)
    ptr = buf.data_ptr() + 4
    assert ptr % 16 != 0
    for fname in ("both", "f2"):
        for probe in (False, True):
            got = run_aggr(ptr, cache, gq, FILTERS[fname], probe, AGGRS[0], data.n, gq.fact_sg)
            assert got == expected(data, FILTERS[fname], probe, AGGRS[0]), (fname, probe)


def test_aggr_scan_compact_date_table_over_16M_rows(cuda):
    """Filters + date probe over 16.4M rows: the fact scan is long enough
    (>= 256 rows per date slot) for the date table's compact form to be
    derived and staged in LDS."""
    from mordred_amd import query as Q
    n = 15 * SEG + 700_000
    q, data = cases.star("q11", n, seed=7)
    print(f"[gpu] aggr scan over {n} rows (compact date table)", flush=True)
    cache = Q.load_cache(data, q, shuffle_seed=11)
    gq = Q.GpuQuery(q, data, cache)
    gq.run()
    sync()
    assert oracle.res_scalar(gq.result()) == expected(data, FILTERS["both"], True, AGGRS[0])
