"""Shared test inputs: golden fixtures and seeded star schemas."""
from __future__ import annotations

import json
import os

import numpy as np

from mordred_amd import query as Q

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

QUERIES = {
    "q11_tiny": lambda: Q.q11(1994),
    "q21_tiny": lambda: Q.q21(60, 500),
    "q4way_tiny": lambda: Q.q4way(60, 500, 300),
    "q21_code0": lambda: Q.q21(40, 300, region=0, category=0),
}


def golden_names():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return [m["name"] for m in json.load(f)]


def load_golden(name):
    """(query, StarData, expected dict) of one fixture."""
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    fact, tables = {}, {}
    for k in z.files:
        if "/" in k:
            t, c = k.split("/", 1)
            (fact if t == "fact" else tables.setdefault(t, {}))[c] = z[k]
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        meta = {m["name"]: m for m in json.load(f)}[name]
    q = QUERIES[name]()
    n = len(next(iter(fact.values())))
    data = Q.StarData(n, meta["seed"], meta["dims"], fact, tables)
    exp = {"scalar": int(z["scalar"][0]),
           "groups": {tuple(int(v) for v in k): int(s) for k, s in zip(z["group_keys"], z["group_sums"])},
           "probe_rows": z["probe_rows"],
           "ht": {int(k[2:]): z[k] for k in z.files if k.startswith("ht")}}
    return q, data, exp


def star(kind: str, n: int, seed: int):
    """A seeded multi-segment star schema of one of the three plain query shapes,
    with small dimensions so tests stay fast."""
    if kind == "q11":
        q = Q.q11(1993)
        return q, Q.make_data(q, n, seed)
    if kind == "q21":
        q = Q.q21(2000, 30000)
        return q, Q.make_data(q, n, seed, S=2000, P=30000)
    q = Q.q4way(2000, 30000, 40000)
    return q, Q.make_data(q, n, seed, S=2000, P=30000, Cn=40000)
