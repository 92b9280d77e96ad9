"""Writes the golden fixtures under tests/golden/ (run from the repo root:
``python tests/golden/gen_golden.py``).

The reference ships no golden vectors, known-answer tests or fixtures for this
path (its test/ directory holds only the SSB data generator and SQL text), so
these fixtures are this project's own.  Inputs are small seeded star schemas;
expected outputs come from a brute-force statement of each query written with
plain Python loops and dictionaries -- independent of the oracle (C), of the
pandas statement in tests/semantics.py and of the HIP kernels.

Each case is one .npz (no pickles) holding the input columns under
"fact/<col>" and "<table>/<col>", plus expected outputs:
  scalar           int64 result of an aggregate query
  group_keys/sums  (G,4) int32 keys and (G,) int64 sums of a group-by query
  probe_rows       (M,5) int32 sorted (lo_row, dim_row1..4) of probe_CPU
                   (rows of tables that are not probed are 0, the CPU value)
  ht<slot>         the hash table the CPU build produces for each dimension
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mordred_amd import query as Q  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def brute_dim(j, table):
    """key -> (dim_row, value) for dimension rows passing the dimension filter."""
    out = {}
    keys = [int(v) for v in table[j.dim_key]]
    vals = [int(v) for v in table[j.group_col]] if j.group_col else [0] * len(keys)
    filt = [int(v) for v in table[j.filter_col]] if j.filter_col else None
    for r, k in enumerate(keys):
        if filt is not None:
            x = filt[r]
            if j.filter_mode == 1 and not (j.c1 <= x <= j.c2):
                continue
            if j.filter_mode == 2 and not (x == j.c1 or x == j.c2):
                continue
            if j.filter_mode == 0 and x != j.c1:
                continue
        out[k] = (r, vals[r])
    return out


def brute_query(q, data):
    dims = {j.slot: brute_dim(j, data.tables[j.table]) for j in q.joins}
    fact = {c: [int(v) for v in a] for c, a in data.fact.items()}
    scalar = 0
    groups = {}
    probe_rows = []
    for i in range(data.n):
        ok = True
        for c, lo, hi in q.fact_filters:
            if not (lo <= fact[c][i] <= hi):
                ok = False
                break
        if not ok:
            continue
        vals = [0, 0, 0, 0]
        drows = [0, 0, 0, 0]
        for j in q.joins:
            hit = dims[j.slot].get(fact[j.fact_key][i])
            if hit is None:
                ok = False
                break
            drows[j.slot - 1], vals[j.slot - 1] = hit
        if not ok:
            continue
        a1 = fact[q.aggr[0]][i]
        a2 = fact[q.aggr[1]][i] if q.aggr[1] else 0
        if q.kind == "aggr":
            scalar += a1 * a2
        else:
            groups[tuple(vals)] = groups.get(tuple(vals), 0) + (a1 - a2)
        probe_rows.append((i, *drows))
    return scalar, groups, sorted(probe_rows)


def brute_ht(j, table):
    ht = [0] * (2 * j.dim_len)
    for k, (r, v) in brute_dim(j, table).items():
        h = (k - j.min_key) % j.dim_len
        ht[2 * h] = v
        ht[2 * h + 1] = r + 1
    return np.array(ht, dtype=np.int32)


CASES = {
    "q11_tiny": (lambda: Q.q11(1994), dict(n=3000, seed=11)),
    "q21_tiny": (lambda: Q.q21(60, 500), dict(n=10000, seed=21, S=60, P=500)),
    "q4way_tiny": (lambda: Q.q4way(60, 500, 300), dict(n=12000, seed=31, S=60, P=500, Cn=300)),
    # brand 0 and region 0: group code 0 must be kept (SSB codes start at 0)
    "q21_code0": (lambda: Q.q21(40, 300, region=0, category=0), dict(n=2000, seed=5, S=40, P=300)),
}


def make_case(name):
    mk, kw = CASES[name]
    q = mk()
    data = Q.make_data(q, **kw)
    scalar, groups, probe_rows = brute_query(q, data)
    arrays = {f"fact/{c}": a for c, a in data.fact.items()}
    for t, cols in data.tables.items():
        for c, a in cols.items():
            arrays[f"{t}/{c}"] = a
    arrays["scalar"] = np.array([scalar], dtype=np.int64)
    keys = sorted(groups)
    arrays["group_keys"] = np.array(keys, dtype=np.int32).reshape(-1, 4)
    arrays["group_sums"] = np.array([groups[k] for k in keys], dtype=np.int64)
    arrays["probe_rows"] = np.array(probe_rows, dtype=np.int32).reshape(-1, 5)
    for j in q.joins:
        arrays[f"ht{j.slot}"] = brute_ht(j, data.tables[j.table])
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrays)
    return {"name": name, "n": data.n, "seed": data.seed, "dims": data.dims, "scalar": scalar,
            "groups": len(groups), "probe_rows": len(probe_rows)}


def main():
    meta = [make_case(n) for n in CASES]
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
