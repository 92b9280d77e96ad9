"""Seeded synthetic star-schema tables (int32 columns, SSB-shaped).

Every column is a counter-based SplitMix64 stream: element i of column c is
output i+1 of SplitMix64 seeded with  mix64(seed ^ (stream(c) * 0xD1B54A32D192ED03)),
mapped to [lo, hi] by the multiply-shift  lo + ((z >> 32) * (hi-lo+1)) >> 32.
Being position-addressable, any row range (a segment, a test slice) can be
generated on its own and always equals the same rows of the full column.

Schema (dictionary-coded like test/ssb/loader/convert.py:48-96):
  date      2,556 rows: d_datekey = yyyymmdd of 1992-01-01 .. 1998-12-30,
            d_year = d_datekey // 10000 (so the date join on d_year equals an
            orderdate range, as the standalone Q1.1 assumes)
  supplier  s_suppkey 1..S, s_nation U{0..24}, s_region = s_nation // 5
  customer  c_custkey 1..C, c_nation U{0..24}, c_region = c_nation // 5
  part      p_partkey 1..P, p_category U{0..24}, p_brand1 U{0..999}
  lineorder lo_orderdate uniform over the date keys, lo_discount U{0..10},
            lo_quantity U{1..50}, lo_extendedprice U{90,000..10,500,000},
            lo_suppkey U{1..S}, lo_partkey U{1..P}, lo_custkey U{1..C},
            lo_revenue U{81,000..10,500,000}, lo_supplycost U{54,000..126,000}
"""
from __future__ import annotations

import datetime as _dt
import zlib

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_STREAM_MUL = 0xD1B54A32D192ED03
_MASK64 = (1 << 64) - 1

LO_RANGES = {
    "lo_discount": (0, 10),
    "lo_quantity": (1, 50),
    "lo_extendedprice": (90_000, 10_500_000),
    "lo_revenue": (81_000, 10_500_000),
    "lo_supplycost": (54_000, 126_000),
}


def _mix64_int(z: int) -> int:
    z &= _MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    return z ^ (z >> 31)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def stream_id(name: str) -> int:
    return zlib.crc32(name.encode())


def stream_key(seed: int, name: str) -> int:
    return _mix64_int(seed ^ ((stream_id(name) * _STREAM_MUL) & _MASK64))


def uniform(seed: int, name: str, lo: int, hi: int, n: int, start: int = 0,
            chunk: int = 1 << 24) -> np.ndarray:
    """Rows [start, start+n) of column `name`: int32 uniform over [lo, hi]."""
    assert 0 < hi - lo + 1 <= 1 << 32
    key = np.uint64(stream_key(seed, name))
    span = np.uint64(hi - lo + 1)
    out = np.empty(n, dtype=np.int32)
    with np.errstate(over="ignore"):
        for c0 in range(0, n, chunk):
            c1 = min(n, c0 + chunk)
            i = np.arange(start + c0 + 1, start + c1 + 1, dtype=np.uint64)
            z = _mix64(key + i * GOLDEN)
            out[c0:c1] = (np.int64(lo) + ((z >> np.uint64(32)) * span >> np.uint64(32)).astype(np.int64))
    return out


def date_keys() -> np.ndarray:
    """The 2,556 SSB date keys 19920101 .. 19981230 as yyyymmdd."""
    d0 = _dt.date(1992, 1, 1)
    days = [d0 + _dt.timedelta(days=i) for i in range(2556)]
    return np.array([d.year * 10000 + d.month * 100 + d.day for d in days], dtype=np.int32)


DATE_MIN_KEY = 19920101
DATE_HT_LEN = 19981230 - 19920101 + 1   # src/gpudb/QueryOptimizer.cu:2719


def make_date() -> dict:
    k = date_keys()
    return {"d_datekey": k, "d_year": (k // 10000).astype(np.int32),
            "d_yearmonthnum": (k // 100).astype(np.int32)}


def make_supplier(n: int, seed: int) -> dict:
    nation = uniform(seed, "s_nation", 0, 24, n)
    return {"s_suppkey": np.arange(1, n + 1, dtype=np.int32), "s_nation": nation,
            "s_region": (nation // 5).astype(np.int32)}


def make_customer(n: int, seed: int) -> dict:
    nation = uniform(seed, "c_nation", 0, 24, n)
    return {"c_custkey": np.arange(1, n + 1, dtype=np.int32), "c_nation": nation,
            "c_region": (nation // 5).astype(np.int32)}


def make_part(n: int, seed: int) -> dict:
    return {"p_partkey": np.arange(1, n + 1, dtype=np.int32),
            "p_category": uniform(seed, "p_category", 0, 24, n),
            "p_brand1": uniform(seed, "p_brand1", 0, 999, n)}


def lineorder_column(name: str, n: int, seed: int, dims: dict, start: int = 0,
                     sorted_rows: int = 0) -> np.ndarray:
    """Rows [start, start+n) of one fact column.  dims: {'S','P','C'} sizes.

    sorted_rows > 0 gives the layout of a lineorder sorted by lo_orderdate
    (test/ssb/sort.c:123 sorts on column 5) of sorted_rows rows: row i holds
    date key i * 2556 // sorted_rows, so every date covers an equal run of rows.
    The other columns do not depend on the order."""
    if name == "lo_orderdate":
        if sorted_rows > 0:
            rows = np.arange(start, start + n, dtype=np.int64)
            return date_keys()[(rows * 2556) // sorted_rows]
        return date_keys()[uniform(seed, name, 0, 2555, n, start)]
    if name == "lo_suppkey":
        return uniform(seed, name, 1, dims["S"], n, start)
    if name == "lo_partkey":
        return uniform(seed, name, 1, dims["P"], n, start)
    if name == "lo_custkey":
        return uniform(seed, name, 1, dims["C"], n, start)
    lo, hi = LO_RANGES[name]
    return uniform(seed, name, lo, hi, n, start)


def make_lineorder(n: int, seed: int, dims: dict, columns, threads: int = 8,
                   sort_orderdate: bool = False) -> dict:
    """Fact columns, generated concurrently (numpy releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    columns = list(columns)
    srt = n if sort_orderdate else 0
    with ThreadPoolExecutor(max(1, min(threads, len(columns)))) as ex:
        arrays = list(ex.map(lambda c: lineorder_column(c, n, seed, dims, sorted_rows=srt), columns))
    return dict(zip(columns, arrays))
