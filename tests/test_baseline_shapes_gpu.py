"""GPU parity at the workload shapes themselves, not reduced ones.

  * filter+join+sum over 6,001,215 fact rows (SSB Q1.1 shape, date dimension
    of 2,556 rows);
  * 3-way star join + group-by over 120,000,000 fact rows with the full
    dimensions: supplier 40,000, part 1,400,000, date 2,556;
  * 4-way star join + group-by with segment caching: full dimensions
    (customer 3,000,000), 64 full segments plus a partial one, half of the
    segments cached by the cost model, the rest on the CPU operators or
    streamed;
  * all 64 queries of the seeded star mix (seed 2024) in one shared-scan
    batch over 120,000,000 fact rows.

Every answer is compared word for word with the oracle (GPU convention) on the
same inputs; the Q1.1 scalar also with its closed form (d_year = datekey /
10000, so the date join is an orderdate range).
"""
from __future__ import annotations

import gc
import os

import numpy as np
import pytest

import oracle
from oracle import queries as OQ
from tests import semantics

pytestmark = pytest.mark.gpu

SEG = oracle.SEGMENT_SIZE
FULL_DIMS = {"S": 40_000, "P": 1_400_000, "C": 3_000_000}


@pytest.fixture(scope="module", autouse=True)
def _oracle_threads():
    oracle.set_threads(min(16, len(os.sched_getaffinity(0))))
    yield
    gc.collect()


def _sync():
    import torch
    torch.cuda.synchronize()


def test_q11_full_size(cuda):
    from mordred_amd import query as Q
    n = 6_001_215
    q = Q.q11(1993)
    data = Q.make_data(q, n, seed=2024)
    print(f"[gpu] Q1.1 shape, {n} rows", flush=True)
    cache = Q.load_cache(data, q, shuffle_seed=21)
    gq = Q.GpuQuery(q, data, cache)
    for _ in range(2):
        gq.run()
        _sync()
        res = gq.result()
        np.testing.assert_array_equal(res, OQ.run_query(q, data, oracle.CONV_GPU))
    f = data.fact
    keep = ((f["lo_discount"] >= 1) & (f["lo_discount"] <= 3) & (f["lo_quantity"] <= 24)
            & (f["lo_orderdate"] // 10000 == 1993))
    closed = int((f["lo_extendedprice"][keep].astype(np.int64) * f["lo_discount"][keep]).sum())
    assert oracle.res_scalar(res) == closed > 0


def test_q21_full_size(cuda):
    from mordred_amd import query as Q
    n = 120_000_000
    q = Q.q21(FULL_DIMS["S"], FULL_DIMS["P"])
    data = Q.make_data(q, n, seed=2024, S=FULL_DIMS["S"], P=FULL_DIMS["P"])
    print(f"[gpu] Q2.1 shape, {n} rows, S=40000 P=1400000", flush=True)
    cache = Q.load_cache(data, q, shuffle_seed=22)
    gq = Q.GpuQuery(q, data, cache)
    gq.run()
    _sync()
    res = gq.result()
    ref = OQ.run_query(q, data, oracle.CONV_GPU)
    np.testing.assert_array_equal(res, ref)
    assert (oracle.res_sums(res) != 0).sum() > 6_000
    del gq, cache, data
    gc.collect()


@pytest.fixture(scope="module")
def q4way_engine(cuda):
    from mordred_amd import engine as E
    from mordred_amd import query as Q
    n = 64 * SEG + 54_321
    q = Q.q4way(FULL_DIMS["S"], FULL_DIMS["P"], FULL_DIMS["C"])
    data = Q.make_data(q, n, seed=2024, S=FULL_DIMS["S"], P=FULL_DIMS["P"], Cn=FULL_DIMS["C"])
    eng = E.Engine(device=0, cache_segments=6 * 65 + 64, threads=min(16, len(os.sched_getaffinity(0))))
    eng.add_data(data, q)
    ref = OQ.run_query(q, data, oracle.CONV_GPU)
    yield q, data, eng, ref
    eng.close()
    del data
    gc.collect()


@pytest.mark.parametrize("placement", ["RESIDENCY", "COST_SPLIT", "GPU_ONLY"])
def test_q4way_segment_cache_full_dims(q4way_engine, placement):
    from mordred_amd import engine as E
    q, data, eng, ref = q4way_engine
    print(f"[gpu] 4-way join, {data.n} rows, 50% cached, {placement}", flush=True)
    eng.place(q, 0.5)
    res, st = eng.run(q, getattr(E, f"PLACE_{placement}"))
    np.testing.assert_array_equal(res, ref)
    assert st["gpu_segments"] == 33          # place(): round(0.5 * 65)
    assert st["gpu_rows"] + st["cpu_rows"] + st["stream_rows"] + st["prefix_rows"] == data.n
    assert semantics.decode_groups(res) == semantics.decode_groups(ref)
    # s_region = 1 keeps the 5 supplier nations 5..9: 5 x 25 x 7 groups
    assert (oracle.res_sums(res) != 0).sum() == 5 * 25 * 7


def test_mix64_batch_every_query(cuda):
    """All 64 queries of the seed-2024 mix over 120M rows in one shared-scan
    batch, each equal to the oracle."""
    from mordred_amd import engine as E
    from mordred_amd import mix
    n = 120_000_000
    qs = [m.query for m in mix.make_mix(64, 2024, FULL_DIMS)]
    data = mix.make_mix_data(qs, n, seed=2024, S=FULL_DIMS["S"], P=FULL_DIMS["P"], Cn=FULL_DIMS["C"])
    cols = mix.fact_columns(qs)
    nseg = (n + SEG - 1) // SEG
    eng = E.Engine(device=0, cache_segments=len(cols) * nseg + 64, threads=min(16, len(os.sched_getaffinity(0))))
    try:
        eng.add_mix_data(data, qs)
        print(f"[gpu] 64-query mix, {n} rows, {len(cols)} columns", flush=True)
        outs, st = eng.batch(qs)
        _sync()
        for i, (q, r) in enumerate(zip(qs, outs)):
            np.testing.assert_array_equal(r, OQ.run_query(q, data, oracle.CONV_GPU), err_msg=f"{i} {q.name}")
        assert st["lds_queries"] + st["global_queries"] + st["dead_queries"] == 64
    finally:
        eng.close()
        del data
        gc.collect()
