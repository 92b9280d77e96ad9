"""On-disk columnar tables in the reference's layout.

One headerless little-endian int32 file per column (common.h:137-201).  The
file of a column is named after its table and its position in the table's
SSB column list (common.h:84-107 lookup): lo_orderdate is LINEORDER5, and
LINEORDERSORT5 in the copy of the table sorted by lo_orderdate (lookupSort,
common.h:109-135; test/ssb/sort.c:123).  The per-segment statistics file of a
column is "<column>minmax" with one "<min> <max>" line per 2^20-row segment
(minmax.cpp:28-45, read by CacheManager.cu:154-184).

Reading and parsing happen in the engine (mordred_engine_load_column_file,
mordred_engine_load_segment_minmax); this module names the files and writes
tables from numpy.
"""
from __future__ import annotations

import os

import numpy as np

from . import abi

# SSB column order of each table: a column's file index is its position here.
TABLE_COLUMNS = {
    "LINEORDER": ("lo_orderkey", "lo_linenumber", "lo_custkey", "lo_partkey", "lo_suppkey", "lo_orderdate",
                  "lo_orderpriority", "lo_shippriority", "lo_quantity", "lo_extendedprice", "lo_ordtotalprice",
                  "lo_discount", "lo_revenue", "lo_supplycost", "lo_tax", "lo_commitdate", "lo_shipmode"),
    "PART": ("p_partkey", "p_name", "p_mfgr", "p_category", "p_brand1", "p_color", "p_type", "p_size",
             "p_container"),
    "SUPPLIER": ("s_suppkey", "s_name", "s_address", "s_city", "s_nation", "s_region", "s_phone"),
    "CUSTOMER": ("c_custkey", "c_name", "c_address", "c_city", "c_nation", "c_region", "c_phone",
                 "c_mktsegment"),
    "DDATE": ("d_datekey", "d_date", "d_dayofweek", "d_month", "d_year", "d_yearmonthnum", "d_yearmonth",
              "d_daynuminweek", "d_daynuminmonth", "d_daynuminyear", "d_sellingseason", "d_lastdayinweekfl",
              "d_lastdayinmonthfl", "d_holidayfl", "d_weekdayfl"),
}
_PREFIX = {"l": "LINEORDER", "p": "PART", "s": "SUPPLIER", "c": "CUSTOMER", "d": "DDATE"}


def file_name(column: str, sorted_table: bool = False) -> str:
    """File name of an SSB column, e.g. lo_orderdate -> LINEORDER5
    (LINEORDERSORT5 with sorted_table)."""
    table = _PREFIX.get(column[:1])
    if table is None or column not in TABLE_COLUMNS[table]:
        raise KeyError(f"not an SSB column: {column}")
    return f"{table}{'SORT' if sorted_table else ''}{TABLE_COLUMNS[table].index(column)}"


def minmax_name(column: str) -> str:
    return column + "minmax"


def write_column(path, values: np.ndarray):
    np.ascontiguousarray(values, dtype="<i4").tofile(path)


def segment_minmax(values: np.ndarray, segment: int = abi.SEGMENT_SIZE):
    """Per-segment (mins, maxs) of a column (the partial last segment counts
    only its own rows)."""
    v = np.asarray(values)
    nseg = (len(v) + segment - 1) // segment
    mins = np.array([v[s * segment:(s + 1) * segment].min() for s in range(nseg)], dtype=np.int64)
    maxs = np.array([v[s * segment:(s + 1) * segment].max() for s in range(nseg)], dtype=np.int64)
    return mins, maxs


def write_minmax(path, mins, maxs):
    with open(path, "w") as f:
        for a, b in zip(mins, maxs):
            f.write(f"{int(a)} {int(b)}\n")


def write_table(data_dir, columns: dict, sorted_table: bool = False, minmax: bool = True) -> dict:
    """Write {column: values} as column files (and minmax files) under
    data_dir; returns {column: path}."""
    os.makedirs(data_dir, exist_ok=True)
    paths = {}
    for name, values in columns.items():
        path = os.path.join(data_dir, file_name(name, sorted_table))
        write_column(path, values)
        if minmax:
            write_minmax(os.path.join(data_dir, minmax_name(name)), *segment_minmax(values))
        paths[name] = path
    return paths


def load_table(engine, data_dir, columns, sorted_table: bool = False, gpu_resident: bool = False,
               minmax: bool = True, n: int = -1):
    """Load column files into an engine; with minmax, also each column's
    minmax file when it exists."""
    for name in columns:
        engine.load_column_file(name, os.path.join(data_dir, file_name(name, sorted_table)), n, gpu_resident)
        mm = os.path.join(data_dir, minmax_name(name))
        if minmax and os.path.exists(mm):
            engine.load_segment_minmax(name, mm)
