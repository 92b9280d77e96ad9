"""Seeded mixes of star queries for shared-scan batches and the serving loop.

A query of the mix joins 1-4 of the dimensions (supplier, customer, part,
date), has 0-3 range predicates and a target selectivity drawn log-uniform
over [1e-4, 0.5], and groups into 1..17,500 dense groups:

  * range predicates (BETWEEN) sit on wide fact columns (lo_revenue,
    lo_extendedprice, lo_supplycost; at most two per query, the reference's
    filter_idx1/2) or on a joined dimension's key (one filter per dimension,
    the build filter of build_GPU2); each gets target**(1/k) of the rows;
  * a query with no range predicate reaches its target through equality
    filters on categorical dimension columns (s_region=r, c_nation=n,
    p_brand1=b, d_year=y, ...), the closest product of their selectivities,
    written as BETWEEN v AND v (mode 1) so both reference paths agree;
  * group columns are chosen per joined dimension from two candidates
    (s_nation/s_region, c_nation/c_region, p_brand1/p_category,
    d_year/d_yearmonthnum) so that the product of their value spans is the
    closest to a target drawn log-uniform over [1, 17,500]; strides form a
    mixed radix, so the dense group index is injective;
  * a one-group query that joins only the date table is an aggregate
    SUM(lo_extendedprice * lo_discount) (the Q1.x form) half of the time,
    otherwise group-by queries sum lo_revenue or lo_revenue - lo_supplycost.
"""
from __future__ import annotations

import dataclasses
import itertools
import math

import numpy as np

from . import datagen
from .query import DimJoin, StarData, StarQuery

# slot -> (table, fact key, dim key, group candidates {col: (min, span)},
#          equality filter candidates {col: (values lo, hi)})
DIMS = {
    1: ("supplier", "lo_suppkey", "s_suppkey",
        {"s_nation": (0, 25), "s_region": (0, 5)}, {"s_region": (0, 4), "s_nation": (0, 24)}),
    2: ("customer", "lo_custkey", "c_custkey",
        {"c_nation": (0, 25), "c_region": (0, 5)}, {"c_region": (0, 4), "c_nation": (0, 24)}),
    3: ("part", "lo_partkey", "p_partkey",
        {"p_brand1": (0, 1000), "p_category": (0, 25)}, {"p_category": (0, 24), "p_brand1": (0, 999)}),
    4: ("date", "lo_orderdate", "d_datekey",
        {"d_year": (1992, 7), "d_yearmonthnum": (199201, 612)}, {"d_year": (1992, 1998)}),
}
FACT_RANGE_COLS = ("lo_revenue", "lo_extendedprice", "lo_supplycost")
MAX_GROUPS = 17_500


@dataclasses.dataclass
class MixQuery:
    query: StarQuery
    target_selectivity: float
    expected_selectivity: float
    target_groups: float
    range_predicates: int


def _dim_len_min(slot: int, dims: dict):
    if slot == 4:
        return datagen.DATE_HT_LEN, datagen.DATE_MIN_KEY
    size = {1: dims["S"], 2: dims["C"], 3: dims["P"]}[slot]
    return size, 0


def _choose_groups(rng, slots, target):
    """Group columns per slot (None = no group value) minimising the log
    distance of the product of spans to target, within MAX_GROUPS."""
    options = [[None] + list(DIMS[s][3]) for s in slots]
    best, best_d = None, None
    for combo in itertools.product(*options):
        g = 1
        for s, c in zip(slots, combo):
            if c is not None:
                g *= DIMS[s][3][c][1]
        if g > MAX_GROUPS:
            continue
        d = abs(math.log(g) - math.log(target)) + 1e-9 * rng.random()
        if best_d is None or d < best_d:
            best, best_d = combo, d
    return dict(zip(slots, best))


def _choose_equalities(rng, slots, target):
    """Equality filters (slot -> column) whose selectivity product is closest to target."""
    options = [[None] + list(DIMS[s][4]) for s in slots]
    best, best_d, best_p = None, None, 1.0
    for combo in itertools.product(*options):
        p = 1.0
        for s, c in zip(slots, combo):
            if c is not None:
                lo, hi = DIMS[s][4][c]
                p /= (hi - lo + 1)
        d = abs(math.log(p) - math.log(target)) + 1e-9 * rng.random()
        if best_d is None or d < best_d:
            best, best_d, best_p = combo, d, p
    return {s: c for s, c in zip(slots, best) if c is not None}, best_p


def make_query(rng, name: str, dims: dict, sel_range=(1e-4, 0.5)) -> MixQuery:
    target = math.exp(rng.uniform(math.log(sel_range[0]), math.log(sel_range[1])))
    nj = int(rng.integers(1, 5))
    slots = sorted(int(s) for s in rng.choice([1, 2, 3, 4], nj, replace=False))
    g_target = math.exp(rng.uniform(0.0, math.log(MAX_GROUPS)))
    groups = _choose_groups(rng, slots, g_target)
    npred = int(rng.integers(0, 4))
    # range predicate sites: up to 2 fact columns, one per joined dimension
    sites = [("fact", c) for c in rng.permutation(FACT_RANGE_COLS)[:2]] + [("dim", s) for s in slots]
    order = rng.permutation(len(sites))[:npred]
    chosen = [sites[i] for i in order]
    fact_filters, dim_filters = [], {}
    expected = 1.0
    if npred:
        f = target ** (1.0 / npred)
        for kind, what in chosen:
            if kind == "fact":
                lo, hi = datagen.LO_RANGES[str(what)]
                span = hi - lo + 1
                w = max(1, int(round(f * span)))
                a = int(rng.integers(lo, hi - w + 2))
                fact_filters.append((str(what), a, a + w - 1))
                expected *= w / span
            else:
                s = int(what)
                if s == 4:
                    keys = datagen.date_keys()
                    w = max(1, int(round(f * len(keys))))
                    i = int(rng.integers(0, len(keys) - w + 1))
                    dim_filters[s] = ("d_datekey", 1, int(keys[i]), int(keys[i + w - 1]))
                    expected *= w / len(keys)
                else:
                    size = _dim_len_min(s, dims)[0]
                    w = max(1, int(round(f * size)))
                    a = int(rng.integers(1, size - w + 2))
                    dim_filters[s] = (DIMS[s][2], 1, a, a + w - 1)
                    expected *= w / size
    else:
        eq, p = _choose_equalities(rng, slots, target)
        for s, col in eq.items():
            lo, hi = DIMS[s][4][col]
            v = int(rng.integers(lo, hi + 1))
            # BETWEEN v AND v: mode 0 (EQ) passes every row on the reference
            # CPU path (CPUProcessing.cu:1189-1200), mode 1 means the same on both
            dim_filters[s] = (col, 1, v, v)
        expected = p
    # dense group strides: mixed radix over the grouped slots
    grouped = [s for s in slots if groups[s] is not None]
    total = 1
    stride = {}
    for s in reversed(grouped):
        stride[s] = total
        total *= DIMS[s][3][groups[s]][1]
    joins = []
    for s in slots:
        table, fkey, dkey, _, _ = DIMS[s]
        dim_len, min_key = _dim_len_min(s, dims)
        j = DimJoin(s, table, fkey, dkey, dim_len, min_key)
        if s in dim_filters:
            j.filter_col, j.filter_mode, j.c1, j.c2 = dim_filters[s]
        if groups[s] is not None:
            j.group_col = groups[s]
            j.min_val = DIMS[s][3][groups[s]][0]
            j.unique_val = stride[s]
        joins.append(j)
    if total == 1 and slots == [4] and rng.random() < 0.5:
        # Q1.x form: the reference's aggregate kernel probes the date table only
        for j in joins:
            j.group_col, j.min_val, j.unique_val = None, 0, 0
        q = StarQuery(name, joins, fact_filters, ("lo_extendedprice", "lo_discount"), "aggr", 1)
    else:
        aggr = ("lo_revenue", None) if rng.random() < 0.5 else ("lo_revenue", "lo_supplycost")
        q = StarQuery(name, joins, fact_filters, aggr, "groupby", total)
    return MixQuery(q, target, expected, g_target, npred)


def make_mix(n_queries: int = 64, seed: int = 7, dims: dict | None = None, sel_range=(1e-4, 0.5)) -> list:
    """n_queries MixQuery, reproducible from the seed."""
    dims = dims or {"S": 40_000, "P": 1_400_000, "C": 3_000_000}
    rng = np.random.default_rng(seed)
    return [make_query(rng, f"mix{seed}_{i}", dims, sel_range) for i in range(n_queries)]


def fact_columns(queries) -> list:
    cols = []
    for q in queries:
        cols += q.fact_columns()
    return list(dict.fromkeys(cols))


def make_mix_data(queries, n: int, seed: int, S: int = 40_000, P: int = 1_400_000,
                  Cn: int = 3_000_000) -> StarData:
    """Every fact and dimension column the queries read."""
    dims = {"S": S, "P": P, "C": Cn}
    tables = {"date": datagen.make_date(), "supplier": datagen.make_supplier(S, seed),
              "customer": datagen.make_customer(Cn, seed), "part": datagen.make_part(P, seed)}
    fact = datagen.make_lineorder(n, seed, dims, fact_columns(queries))
    return StarData(n, seed, dims, fact, tables)
