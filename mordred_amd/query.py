"""Star-query plans and their execution through the C ABI.

A :class:`StarQuery` is the information the reference optimizer hands to the
operator layer for one star query (src/gpudb/QueryOptimizer.cu prepareQuery):
per joined dimension its hash-table slot (table_id-1: supplier 1, customer 2,
part 3, date 4; CacheManager.cu:1243-1271), key range, dimension filter and
group column, plus the fact filters, aggregate columns and dense group-index
strides.  :func:`run_gpu` performs what CPUGPUProcessing::call_bfilter_build_GPU
and call_{pfilter_}probe_{group_by,aggr}_GPU do for a fully cached segment
group: zero the hash tables and result, build every dimension table, launch the
fused fact kernel, all on one stream.

Device memory is plain torch tensors (a segment cache is one int32 tensor);
all arithmetic runs in libmordred_hip.so.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional

import numpy as np

from . import abi, datagen

SEG = abi.SEGMENT_SIZE


@dataclasses.dataclass
class DimJoin:
    slot: int                    # 1..4: supplier, customer, part, date
    table: str                   # dimension table name ("supplier", ...)
    fact_key: str                # fact foreign-key column
    dim_key: str                 # dimension primary-key column
    dim_len: int                 # hash-table slots (KernelArgs probeArgs dim_len)
    min_key: int                 # HASH minimum
    filter_col: Optional[str] = None
    filter_mode: int = 1         # 0 EQ, 1 BETWEEN, 2 EQ-or-EQ (build_GPU2 modes)
    c1: int = 0
    c2: int = 0
    group_col: Optional[str] = None   # value stored in the hash table
    min_val: int = 0
    unique_val: int = 0


@dataclasses.dataclass
class StarQuery:
    name: str
    joins: list
    fact_filters: list           # [(column, lo, hi)] BETWEEN, at most 2
    aggr: tuple                  # (col1, col2 or None)
    kind: str                    # "aggr" (sum a1*a2) or "groupby" (sum a1-a2 per group)
    total_val: int = 1

    def fact_columns(self) -> list:
        cols = [c for c, _, _ in self.fact_filters] + [j.fact_key for j in self.joins]
        cols += [a for a in self.aggr if a]
        return list(dict.fromkeys(cols))

    def join(self, slot: int) -> Optional[DimJoin]:
        for j in self.joins:
            if j.slot == slot:
                return j
        return None


# ---------------------------------------------------------------- plans

def q11(year: int = 1993) -> StarQuery:
    """SSB Q1.1 (QueryOptimizer.cu:176-236, 2558-2606, 2708-2750)."""
    date = DimJoin(4, "date", "lo_orderdate", "d_datekey", datagen.DATE_HT_LEN, datagen.DATE_MIN_KEY,
                   filter_col="d_year", filter_mode=1, c1=year, c2=year)
    return StarQuery("q11", [date], [("lo_discount", 1, 3), ("lo_quantity", 0, 24)],
                     ("lo_extendedprice", "lo_discount"), "aggr", 1)


def q12(yearmonth: int = 199401) -> StarQuery:
    """SSB Q1.2 (QueryOptimizer.cu:2607-2656): d_yearmonthnum = 199401,
    lo_discount 4..6, lo_quantity 26..35."""
    date = DimJoin(4, "date", "lo_orderdate", "d_datekey", datagen.DATE_HT_LEN, datagen.DATE_MIN_KEY,
                   filter_col="d_yearmonthnum", filter_mode=1, c1=yearmonth, c2=yearmonth)
    return StarQuery("q12", [date], [("lo_discount", 4, 6), ("lo_quantity", 26, 35)],
                     ("lo_extendedprice", "lo_discount"), "aggr", 1)


def q13(first: int = 19940204, last: int = 19940210) -> StarQuery:
    """SSB Q1.3 (QueryOptimizer.cu:2657-2705): the reference expresses week 6
    of 1994 as d_datekey BETWEEN 19940204 AND 19940210; lo_discount 5..7,
    lo_quantity 26..35."""
    date = DimJoin(4, "date", "lo_orderdate", "d_datekey", datagen.DATE_HT_LEN, datagen.DATE_MIN_KEY,
                   filter_col="d_datekey", filter_mode=1, c1=first, c2=last)
    return StarQuery("q13", [date], [("lo_discount", 5, 7), ("lo_quantity", 26, 35)],
                     ("lo_extendedprice", "lo_discount"), "aggr", 1)


def q21(S: int, P: int, region: int = 1, category: int = 1) -> StarQuery:
    """SSB Q2.1 (QueryOptimizer.cu:362-446, 2755-2800, 2906-2959) with
    p_brand1 over 1,000 codes: group index brand*7 + (year-1992)."""
    supp = DimJoin(1, "supplier", "lo_suppkey", "s_suppkey", S, 0, filter_col="s_region", c1=region, c2=region)
    part = DimJoin(3, "part", "lo_partkey", "p_partkey", P, 0, filter_col="p_category", c1=category,
                   c2=category, group_col="p_brand1", min_val=0, unique_val=7)
    date = DimJoin(4, "date", "lo_orderdate", "d_datekey", datagen.DATE_HT_LEN, datagen.DATE_MIN_KEY,
                   group_col="d_year", min_val=1992, unique_val=1)
    return StarQuery("q21", [supp, part, date], [], ("lo_revenue", None), "groupby", 7 * 1000)


def q4way(S: int, P: int, Cn: int, region: int = 1, category: int = 1) -> StarQuery:
    """4-way join grouped by (c_nation, s_nation, year) = 25x25x7 with the
    Q4.x aggregate revenue - supplycost (strides of Q3.1, QueryOptimizer.cu:3009-3014)."""
    supp = DimJoin(1, "supplier", "lo_suppkey", "s_suppkey", S, 0, filter_col="s_region", c1=region,
                   c2=region, group_col="s_nation", min_val=0, unique_val=25 * 7)
    cust = DimJoin(2, "customer", "lo_custkey", "c_custkey", Cn, 0, group_col="c_nation", min_val=0,
                   unique_val=7)
    part = DimJoin(3, "part", "lo_partkey", "p_partkey", P, 0, filter_col="p_category", c1=category,
                   c2=category)
    date = DimJoin(4, "date", "lo_orderdate", "d_datekey", datagen.DATE_HT_LEN, datagen.DATE_MIN_KEY,
                   group_col="d_year", min_val=1992, unique_val=1)
    return StarQuery("q4way", [supp, cust, part, date], [], ("lo_revenue", "lo_supplycost"), "groupby",
                     7 * 25 * 25)


# ---------------------------------------------------------------- data

@dataclasses.dataclass
class StarData:
    n: int
    seed: int
    dims: dict                  # {"S","P","C"}
    fact: dict                  # column -> int32 numpy array
    tables: dict                # "date"/"supplier"/... -> {column: array}


def make_data(query: StarQuery, n: int, seed: int, S: int = 40_000, P: int = 1_400_000,
              Cn: int = 3_000_000, sort_orderdate: bool = False) -> StarData:
    dims = {"S": S, "P": P, "C": Cn}
    tables = {"date": datagen.make_date()}
    names = {j.table for j in query.joins}
    if "supplier" in names:
        tables["supplier"] = datagen.make_supplier(S, seed)
    if "part" in names:
        tables["part"] = datagen.make_part(P, seed)
    if "customer" in names:
        tables["customer"] = datagen.make_customer(Cn, seed)
    fact = datagen.make_lineorder(n, seed, dims, query.fact_columns(), sort_orderdate=sort_orderdate)
    return StarData(n, seed, dims, fact, tables)


# ---------------------------------------------------------------- device side

def _torch():
    import torch
    return torch


class DeviceSegmentCache:
    """HBM segment cache: one int32 tensor of `capacity` segments of 2^20 values
    (the reference gpuCache, CacheManager.cu:44) plus one device segment table
    per column (col_idx, CacheManager.cu:246-257).  `shuffle` places segments
    at a seeded random permutation of slots so that indirection is exercised."""

    def __init__(self, capacity: int, device: str = "cuda", shuffle_seed: Optional[int] = None):
        torch = _torch()
        self.capacity = capacity
        self.device = device
        self.buf = torch.zeros(capacity * SEG, dtype=torch.int32, device=device)
        order = np.arange(capacity)
        if shuffle_seed is not None:
            np.random.default_rng(shuffle_seed).shuffle(order)
        self.free = list(order[::-1])
        self.tables = {}
        self.lengths = {}

    def load(self, name: str, values: np.ndarray) -> None:
        torch = _torch()
        values = np.ascontiguousarray(values, dtype=np.int32)
        nseg = max(1, (len(values) + SEG - 1) // SEG)
        if nseg > len(self.free):
            raise MemoryError(f"segment cache full loading {name}")
        slots = [int(self.free.pop()) for _ in range(nseg)]
        src = torch.from_numpy(values).to(self.device)
        for s, slot in enumerate(slots):
            part = src[s * SEG:(s + 1) * SEG]
            self.buf[slot * SEG: slot * SEG + part.numel()].copy_(part)
        self.tables[name] = torch.tensor(slots, dtype=torch.int32, device=self.device)
        self.lengths[name] = len(values)

    def idx(self, name: Optional[str]) -> Optional[int]:
        return None if name is None else self.tables[name].data_ptr()

    @property
    def ptr(self) -> int:
        return self.buf.data_ptr()


def segments_needed(data: StarData, query: StarQuery) -> int:
    n = 0
    for c in query.fact_columns():
        n += (data.n + SEG - 1) // SEG
    for j in query.joins:
        t = data.tables[j.table]
        m = len(t[j.dim_key])
        cols = {j.dim_key, j.filter_col, j.group_col} - {None}
        n += len(cols) * ((m + SEG - 1) // SEG)
    return n


def load_cache(data: StarData, query: StarQuery, device: str = "cuda",
               shuffle_seed: Optional[int] = None) -> DeviceSegmentCache:
    cache = DeviceSegmentCache(segments_needed(data, query), device, shuffle_seed)
    for c in query.fact_columns():
        cache.load(c, data.fact[c])
    for j in query.joins:
        t = data.tables[j.table]
        for c in {j.dim_key, j.filter_col, j.group_col} - {None}:
            if c not in cache.tables:
                cache.load(c, t[c])
    return cache


def seg_group(n_rows: int, device: str = "cuda", segments=None):
    """Device segment group covering rows [0, n_rows) (or the given segments)."""
    torch = _torch()
    if segments is None:
        segments = range((n_rows + SEG - 1) // SEG)
    return torch.tensor(list(segments), dtype=torch.int16, device=device)


class GpuQuery:
    """Device state of one query: hash tables, result buffer and the kernel
    argument structs, assembled once like CPUGPUProcessing.cu:233-297."""

    def __init__(self, query: StarQuery, data: StarData, cache: DeviceSegmentCache, stream=None):
        torch = _torch()
        self.q, self.data, self.cache = query, data, cache
        dev = cache.device
        self.stream = stream
        # the result and the hash tables share one buffer, so one fill clears
        # them all; each part starts on a 512-byte boundary, as a separate
        # allocation would (the result first: its atomics then share no cache
        # line with a hash table, which made the Q1.1 kernel 20% slower)
        def pad(n: int) -> int:
            return (n + 127) // 128 * 128
        nres = max(1, query.total_val) * 6
        offs, total = {}, pad(nres)
        for j in query.joins:
            offs[j.slot] = total
            total += pad(2 * j.dim_len)
        self._htbuf = torch.zeros(total, dtype=torch.int32, device=dev)
        self.hts = {j.slot: self._htbuf[offs[j.slot]:offs[j.slot] + 2 * j.dim_len] for j in query.joins}
        self.res = self._htbuf[:nres]
        self.dim_sg = {j.slot: seg_group(len(data.tables[j.table][j.dim_key]), dev) for j in query.joins}
        self.fact_sg = seg_group(data.n, dev)
        p = abi.ProbeArgsGPU()
        g = abi.GroupbyArgsGPU()
        for j in query.joins:
            setattr(p, f"key_idx{j.slot}", cache.idx(j.fact_key))
            setattr(p, f"ht{j.slot}", self.hts[j.slot].data_ptr())
            setattr(p, f"dim_len{j.slot}", j.dim_len)
            setattr(p, f"min_key{j.slot}", j.min_key)
            if j.group_col:
                setattr(g, f"group_idx{j.slot}", cache.idx(j.group_col))
                setattr(g, f"min_val{j.slot}", j.min_val)
                setattr(g, f"unique_val{j.slot}", j.unique_val)
        g.aggr_idx1 = cache.idx(query.aggr[0])
        g.aggr_idx2 = cache.idx(query.aggr[1])
        g.total_val = query.total_val
        f = None
        if query.fact_filters:
            f = abi.FilterArgsGPU()
            (c1, lo1, hi1) = query.fact_filters[0]
            f.filter_idx1, f.compare1, f.compare2, f.mode1 = cache.idx(c1), lo1, hi1, 1
            if len(query.fact_filters) > 1:
                (c2, lo2, hi2) = query.fact_filters[1]
                f.filter_idx2, f.compare3, f.compare4, f.mode2 = cache.idx(c2), lo2, hi2, 1
        self.pargs, self.gargs, self.fargs = p, g, f
        self.bargs = {}
        self.bfargs = {}
        for j in query.joins:
            b = abi.BuildArgsGPU(cache.idx(j.dim_key), cache.idx(j.group_col), j.dim_len, j.min_key, 0)
            bf = None
            if j.filter_col:
                bf = abi.FilterArgsGPU()
                bf.filter_idx1, bf.compare1, bf.compare2, bf.mode1 = cache.idx(j.filter_col), j.c1, j.c2, j.filter_mode
            self.bargs[j.slot], self.bfargs[j.slot] = b, bf

    def _stream(self):
        if self.stream is not None:
            return self.stream
        return _torch().cuda.current_stream().cuda_stream

    def reset(self) -> None:
        self._htbuf.zero_()

    def build(self) -> None:
        s = self._stream()
        for j in self.q.joins:
            n = len(self.data.tables[j.table][j.dim_key])
            abi.call("mordred_build_GPU2", self.cache.ptr, abi.ref(self.bfargs[j.slot]), abi.ref(self.bargs[j.slot]),
                     n, self.hts[j.slot].data_ptr(), 0, self.dim_sg[j.slot].data_ptr(), s)

    def probe(self, n: Optional[int] = None, sg=None) -> None:
        s = self._stream()
        n = self.data.n if n is None else n
        sgp = (self.fact_sg if sg is None else sg).data_ptr()
        if self.q.kind == "aggr":
            abi.call("mordred_filter_probe_aggr_GPU2", self.cache.ptr, abi.ref(self.fargs), C.pointer(self.pargs),
                     C.pointer(self.gargs), n, self.res.data_ptr(), 0, sgp, s)
        else:
            abi.call("mordred_filter_probe_group_by_GPU2", self.cache.ptr, abi.ref(self.fargs),
                     C.pointer(self.pargs), C.pointer(self.gargs), n, self.res.data_ptr(), 0, sgp, s)

    def run(self) -> None:
        self.reset()
        self.build()
        self.probe()

    def result(self) -> np.ndarray:
        return self.res.cpu().numpy()


def run_gpu(query: StarQuery, data: StarData, device: str = "cuda", shuffle_seed: Optional[int] = None):
    cache = load_cache(data, query, device, shuffle_seed)
    gq = GpuQuery(query, data, cache)
    gq.run()
    _torch().cuda.synchronize()
    return gq.result()
