"""ORACLE -- test infrastructure, never the product.

Runs a ``mordred_amd.query.StarQuery`` through the C restatement of the
reference CPU path, the way QueryProcessing::runQuery drives the CPU operators
for a segment group held entirely on the CPU: zeroed hash tables
(QueryOptimizer.cu:2946-2953), one build_CPU per dimension
(CPUGPUProcessing::call_bfilter_build_CPU), then the fused fact operator
(filter_probe_aggr_CPU for Q1.x, {filter_}probe_group_by_CPU otherwise).
"""
from __future__ import annotations

import numpy as np

from . import (CONV_CPU, BuildArgs, FilterArgs, GroupbyArgs, ProbeArgs, SEGMENT_SIZE, iptr, lib,
               sptr)


def segment_group(n_rows: int, segments=None) -> np.ndarray:
    if segments is None:
        segments = range((n_rows + SEGMENT_SIZE - 1) // SEGMENT_SIZE)
    return np.array(list(segments), dtype=np.int16)


def build_tables(query, data, conv: int = CONV_CPU) -> dict:
    """slot -> hash table (int32, 2*dim_len) built by oracle_build_CPU."""
    hts = {}
    for j in query.joins:
        t = data.tables[j.table]
        n = len(t[j.dim_key])
        ht = np.zeros(2 * j.dim_len, dtype=np.int32)
        f = FilterArgs()
        if j.filter_col:
            f.filter_col1, f.compare1, f.compare2, f.mode1 = iptr(t[j.filter_col]), j.c1, j.c2, j.filter_mode
        b = BuildArgs(iptr(t[j.dim_key]), iptr(t[j.group_col]) if j.group_col else None, j.dim_len, j.min_key, 0)
        lib.oracle_build_CPU(f, b, n, iptr(ht), 0, sptr(segment_group(n)), conv)
        hts[j.slot] = ht
    return hts


def fact_args(query, data, hts):
    p, g, f = ProbeArgs(), GroupbyArgs(), FilterArgs()
    keep = []
    for j in query.joins:
        setattr(p, f"key_col{j.slot}", iptr(data.fact[j.fact_key]))
        setattr(p, f"ht{j.slot}", iptr(hts[j.slot]))
        setattr(p, f"dim_len{j.slot}", j.dim_len)
        setattr(p, f"min_key{j.slot}", j.min_key)
        if j.group_col:
            col = np.ascontiguousarray(data.tables[j.table][j.group_col])
            keep.append(col)
            setattr(g, f"group_col{j.slot}", iptr(col))
            setattr(g, f"min_val{j.slot}", j.min_val)
            setattr(g, f"unique_val{j.slot}", j.unique_val)
    g.aggr_col1 = iptr(data.fact[query.aggr[0]])
    g.aggr_col2 = iptr(data.fact[query.aggr[1]]) if query.aggr[1] else None
    g.total_val = query.total_val
    if query.fact_filters:
        c, lo, hi = query.fact_filters[0]
        f.filter_col1, f.compare1, f.compare2, f.mode1 = iptr(data.fact[c]), lo, hi, 1
        if len(query.fact_filters) > 1:
            c, lo, hi = query.fact_filters[1]
            f.filter_col2, f.compare3, f.compare4, f.mode2 = iptr(data.fact[c]), lo, hi, 1
    return f, p, g, keep


def run_query(query, data, conv: int = CONV_CPU, segments=None, n=None) -> np.ndarray:
    """Result buffer (total_val*6 int32) of the whole query on the oracle."""
    hts = build_tables(query, data, conv)
    return run_fact(query, data, hts, conv, segments, n)


def run_fact(query, data, hts, conv: int = CONV_CPU, segments=None, n=None) -> np.ndarray:
    f, p, g, keep = fact_args(query, data, hts)
    res = np.zeros(max(1, query.total_val) * 6, dtype=np.int32)
    sg = segment_group(data.n, segments)
    if n is None:
        n = data.n if segments is None else min(data.n, len(sg) * SEGMENT_SIZE)
    if query.kind == "aggr":
        lib.oracle_filter_probe_aggr_CPU(f, p, g, n, iptr(res), 0, sptr(sg), conv)
    elif query.fact_filters:
        lib.oracle_filter_probe_group_by_CPU(f, p, g, n, iptr(res), 0, sptr(sg), conv)
    else:
        lib.oracle_probe_group_by_CPU(p, g, n, iptr(res), 0, sptr(sg), conv)
    del keep
    return res
