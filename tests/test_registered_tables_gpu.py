"""GPU parity with registered probe tables (mordred_probe_table_register).

A registered hash table's compact forms are derived by the first fused call
that compacts it and reused by the later ones, on any stream, until a build
into the table (or mordred_probe_table_invalidate) marks them stale.  Every
answer must equal the oracle's (GPU convention) for the table contents at
the time of the call:
  * segment groups of one query run on three streams into one result buffer,
    twice (the second pass reuses the forms);
  * row probes (probe_GPU2) after value probes use the other form;
  * a rebuild of the same tables with other dimension filters (a build marks
    the forms stale);
  * tables overwritten by a plain device copy, then invalidated;
  * the no-alloc entry point with registered tables;
  * unregistering restores per-call derivation.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from oracle import CONV_GPU, Offset, iptr, lib, sptr
from oracle import queries as OQ

pytestmark = pytest.mark.gpu

SEG = oracle.SEGMENT_SIZE


def _t():
    import torch
    return torch


def sync():
    _t().cuda.synchronize()


@pytest.fixture(scope="module")
def big(cuda):
    """Part table above the compaction threshold (300,000 slots) and enough
    fact rows that the fused calls compact the supplier and part tables."""
    from mordred_amd import query as Q
    q = Q.q21(5_000, 300_000)
    data = Q.make_data(q, 6 * SEG + 4321, seed=91, S=5_000, P=300_000)
    cache = Q.load_cache(data, q, shuffle_seed=13)
    return q, data, cache


def register(gq):
    from mordred_amd import abi
    for j in gq.q.joins:
        abi.call("mordred_probe_table_register", gq.hts[j.slot].data_ptr(), j.dim_len)


def unregister(gq):
    from mordred_amd import abi
    for j in gq.q.joins:
        abi.call("mordred_probe_table_unregister", gq.hts[j.slot].data_ptr())


def probe_on_streams(gq, data, groups):
    """The fused group-by over each segment group on its own stream, all into
    gq.res; the groups cover the fact table once."""
    from mordred_amd import abi
    from mordred_amd import query as Q
    torch = _t()
    main = torch.cuda.current_stream()
    streams = [torch.cuda.Stream() for _ in groups]
    sgs = [Q.seg_group(0, segments=g) for g in groups]
    for s, sg, g in zip(streams, sgs, groups):
        s.wait_stream(main)
        n = sum(min(SEG, data.n - seg * SEG) for seg in g)
        abi.call("mordred_filter_probe_group_by_GPU2", gq.cache.ptr, abi.ref(gq.fargs), C.pointer(gq.pargs),
                 C.pointer(gq.gargs), n, gq.res.data_ptr(), 0, sg.data_ptr(), s.cuda_stream)
    for s in streams:
        main.wait_stream(s)
    sync()
    return gq.res.cpu().numpy()


def test_registered_tables_across_streams_rebuilds_and_invalidation(big):
    from mordred_amd import abi
    from mordred_amd import query as Q
    q, data, cache = big
    groups = [[0, 3], [1, 4], [2, 5, 6]]      # segment 6 is the short one, last in its group
    gq = Q.GpuQuery(q, data, cache)
    register(gq)
    try:
        ref = OQ.run_query(q, data, CONV_GPU)
        for rep in range(2):
            print(f"[gpu] registered tables: three streams, pass {rep}", flush=True)
            gq.reset()
            gq.build()
            np.testing.assert_array_equal(probe_on_streams(gq, data, groups), ref)

        # row probes of the same tables: the other compact form
        print("[gpu] registered tables: probe_GPU2 rows", flush=True)
        n = data.n
        outs = [_t().zeros(n, dtype=_t().int32, device="cuda") for _ in range(5)]
        total = _t().zeros(1, dtype=_t().int32, device="cuda")
        abi.call("mordred_probe_GPU2", cache.ptr, C.pointer(gq.pargs),
                 C.pointer(abi.OffsetGPU(*[t.data_ptr() for t in outs])), n, total.data_ptr(), 0,
                 gq.fact_sg.data_ptr(), _t().cuda.current_stream().cuda_stream)
        sync()
        hts = OQ.build_tables(q, data, CONV_GPU)
        f, p, g, keep = OQ.fact_args(q, data, hts)
        ro = [np.zeros(n, dtype=np.int32) for _ in range(5)]
        rt = np.zeros(1, dtype=np.int32)
        lib.oracle_probe_CPU(p, Offset(*[iptr(a) for a in ro]), n, iptr(rt), 0, sptr(OQ.segment_group(n)), CONV_GPU)
        m = int(total.item())
        assert m == int(rt[0]) > 0
        got = np.stack([t.cpu().numpy()[:m] for t in outs], axis=1)
        exp = np.stack([a[:m] for a in ro], axis=1)
        np.testing.assert_array_equal(got[np.lexsort(got.T[::-1])], exp[np.lexsort(exp.T[::-1])])

        # rebuild the same tables with other dimension filters: the builds
        # mark the registered forms stale
        print("[gpu] registered tables: rebuilt with region 3, category 7", flush=True)
        q2 = Q.q21(5_000, 300_000, region=3, category=7)
        saved = {}
        for j, j2 in zip(q.joins, q2.joins):
            bf = gq.bfargs[j.slot]
            if bf is not None:
                saved[j.slot] = (bf.compare1, bf.compare2)
                bf.compare1, bf.compare2 = j2.c1, j2.c2
        gq.reset()
        gq.build()
        np.testing.assert_array_equal(probe_on_streams(gq, data, groups), OQ.run_query(q2, data, CONV_GPU))
        for slot, (c1, c2) in saved.items():
            gq.bfargs[slot].compare1, gq.bfargs[slot].compare2 = c1, c2

        # overwrite the tables with a plain copy of the original query's
        # tables, then invalidate: the forms follow the new contents
        print("[gpu] registered tables: copied tables + invalidate", flush=True)
        other = Q.GpuQuery(q, data, cache)
        other.run()
        sync()
        gq.res.zero_()
        for j in q.joins:
            gq.hts[j.slot].copy_(other.hts[j.slot])
            abi.call("mordred_probe_table_invalidate", gq.hts[j.slot].data_ptr())
        np.testing.assert_array_equal(probe_on_streams(gq, data, groups), ref)

        # the no-alloc entry point serves registered tables from their forms
        print("[gpu] registered tables: no-alloc entry point", flush=True)
        gq.res.zero_()
        abi.call("mordred_filter_probe_group_by_GPU2_noalloc", cache.ptr, abi.ref(gq.fargs), C.pointer(gq.pargs),
                 C.pointer(gq.gargs), data.n, gq.res.data_ptr(), 0, gq.fact_sg.data_ptr(),
                 _t().cuda.current_stream().cuda_stream)
        sync()
        np.testing.assert_array_equal(gq.result(), ref)
    finally:
        unregister(gq)
    # unregistered again: per-call derivation, same answers
    print("[gpu] registered tables: after unregister", flush=True)
    gq.reset()
    gq.build()
    np.testing.assert_array_equal(probe_on_streams(gq, data, groups), ref)


def test_register_rejects_bad_arguments(cuda):
    from mordred_amd import abi
    lib_ = abi.load()
    assert lib_.mordred_probe_table_register(None, 10) == 1
    buf = _t().zeros(16, dtype=_t().int32, device="cuda")
    assert lib_.mordred_probe_table_register(buf.data_ptr(), 0) == 1
    assert lib_.mordred_probe_table_unregister(buf.data_ptr()) == 0      # not registered: nothing to do
    assert lib_.mordred_probe_table_invalidate(None) == 1
