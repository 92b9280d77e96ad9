"""The query mix generator, and the oracle / engine CPU operators on mix
queries (pinned by the pandas statement of each query)."""
from __future__ import annotations

import collections

import numpy as np
import pytest

import oracle
from oracle import queries as OQ
from mordred_amd import engine as E
from mordred_amd import mix
from tests import semantics

SMALL = {"S": 2000, "P": 30000, "C": 40000}


def test_mix_is_reproducible():
    a = [m.query for m in mix.make_mix(64, 3, SMALL)]
    b = [m.query for m in mix.make_mix(64, 3, SMALL)]
    assert a == b
    assert [m.query for m in mix.make_mix(64, 4, SMALL)] != a


def test_mix_shape_bounds():
    ms = mix.make_mix(256, 5)
    joins = collections.Counter(len(m.query.joins) for m in ms)
    assert set(joins) == {1, 2, 3, 4}
    for m in ms:
        q = m.query
        assert 1 <= len(q.joins) <= 4 and len(q.fact_filters) <= 2
        assert 0 <= m.range_predicates <= 3
        assert len(q.fact_filters) + sum(1 for j in q.joins if j.filter_col) >= min(1, m.range_predicates)
        assert all(j.filter_mode == 1 for j in q.joins if j.filter_col)
        assert 1 <= q.total_val <= mix.MAX_GROUPS
        assert 1e-4 <= m.target_selectivity <= 0.5
        # mixed-radix strides: the dense index is injective and in range
        spans = {j.slot: mix.DIMS[j.slot][3][j.group_col][1] for j in q.joins if j.group_col}
        prod = int(np.prod(list(spans.values()))) if spans else 1
        assert prod == q.total_val
        for j in q.joins:
            if j.group_col:
                later = [spans[s] for s in spans if s > j.slot]
                assert j.unique_val == (int(np.prod(later)) if later else 1)
        if q.kind == "aggr":
            assert [j.slot for j in q.joins] == [4] and q.total_val == 1
    assert len(mix.fact_columns([m.query for m in ms])) <= 12
    # realised selectivities follow the targets where range predicates set them
    ranged = [m for m in ms if m.range_predicates]
    ratio = np.array([m.expected_selectivity / m.target_selectivity for m in ranged])
    assert np.median(np.abs(np.log(ratio))) < 0.05


@pytest.fixture(scope="module")
def small_mix():
    qs = [m.query for m in mix.make_mix(16, 21, SMALL)]
    data = mix.make_mix_data(qs, 150_000, seed=21, S=SMALL["S"], P=SMALL["P"], Cn=SMALL["C"])
    return qs, data


def test_oracle_matches_pandas_on_mix(small_mix):
    qs, data = small_mix
    for q in qs:
        res = OQ.run_query(q, data, oracle.CONV_GPU)
        if q.kind == "aggr":
            assert oracle.res_scalar(res) == semantics.scalar_sum(q, data), q.name
        else:
            got = semantics.decode_groups(res)
            exp = semantics.group_map(q, data)
            # CONV_GPU writes every key word, so a group keyed all-zero with a
            # zero sum is invisible in both statements alike
            assert got == {k: v for k, v in exp.items() if any(k) or v}, q.name


def test_engine_cpu_operators_on_mix(small_mix):
    qs, data = small_mix
    eng = E.Engine(device=-1, threads=4)
    eng.add_mix_data(data, qs)
    for q in qs:
        res, _ = eng.run(q, E.PLACE_CPU_ONLY)
        np.testing.assert_array_equal(res, OQ.run_query(q, data, oracle.CONV_CPU), err_msg=q.name)
    eng.close()
