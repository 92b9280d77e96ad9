"""Independent statements of the star queries, used to pin the oracle.

These evaluate the SQL meaning of a StarQuery with numpy/pandas joins and
group-bys -- no perfect hashing, no dense group index, no segment addressing --
so a mistake in the oracle's hash slots, probe words, index strides, filters or
accumulation shows up as a mismatch.  Result buffers are decoded by the keys
stored next to each sum, which does not depend on the index formula.
"""
from __future__ import annotations

import numpy as np
import pandas as pd


def dim_frame(query, data, j):
    t = data.tables[j.table]
    df = pd.DataFrame({"key": t[j.dim_key], "row": np.arange(len(t[j.dim_key]), dtype=np.int64)})
    if j.group_col:
        df["val"] = t[j.group_col].astype(np.int64)
    else:
        df["val"] = 0
    if j.filter_col:
        x = t[j.filter_col]
        if j.filter_mode == 1:
            keep = (x >= j.c1) & (x <= j.c2)
        elif j.filter_mode == 2:
            keep = (x == j.c1) | (x == j.c2)
        else:
            keep = x == j.c1
        df = df[keep]
    return df


def joined(query, data, rows=None):
    """Fact rows (optionally restricted to `rows`) that pass every fact filter
    and join every dimension; columns g<slot> hold each dimension's value."""
    idx = np.arange(data.n, dtype=np.int64) if rows is None else np.asarray(rows, dtype=np.int64)
    df = pd.DataFrame({"lo_row": idx})
    for c, lo, hi in query.fact_filters:
        x = data.fact[c][idx]
        df = df[(x >= lo) & (x <= hi)]
        idx = df["lo_row"].to_numpy()
    for j in query.joins:
        df[j.fact_key + "_k"] = data.fact[j.fact_key][df["lo_row"].to_numpy()]
        d = dim_frame(query, data, j).rename(columns={"key": j.fact_key + "_k", "row": f"d{j.slot}",
                                                       "val": f"g{j.slot}"})
        df = df.merge(d, on=j.fact_key + "_k", how="inner")
    a1 = data.fact[query.aggr[0]][df["lo_row"].to_numpy()].astype(np.int64)
    a2 = (data.fact[query.aggr[1]][df["lo_row"].to_numpy()].astype(np.int64) if query.aggr[1]
          else np.zeros(len(df), dtype=np.int64))
    df["a1"], df["a2"] = a1, a2
    return df


def scalar_sum(query, data, rows=None) -> int:
    df = joined(query, data, rows)
    return int((df["a1"] * df["a2"]).sum())


def group_map(query, data, rows=None) -> dict:
    """{(v1, v2, v3, v4): int64 sum of a1-a2}; v_k = 0 for a table without a value."""
    df = joined(query, data, rows)
    for k in range(1, 5):
        if f"g{k}" not in df:
            df[f"g{k}"] = 0
    df["d"] = df["a1"] - df["a2"]
    out = df.groupby(["g1", "g2", "g3", "g4"])["d"].sum()
    return {tuple(int(v) for v in k): int(s) for k, s in out.items()}


def decode_groups(res: np.ndarray) -> dict:
    """{keys: sum} over result groups that hold any non-zero word."""
    r = res.reshape(-1, 6)
    sums = r[:, 4:6].copy().view(np.int64).reshape(-1)
    out = {}
    for h in np.nonzero(np.any(r[:, :4] != 0, axis=1) | (sums != 0))[0]:
        out[tuple(int(v) for v in r[h, :4])] = int(sums[h])
    return out
