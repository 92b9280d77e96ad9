"""The no-alloc entry point, mordred_filter_probe_group_by_GPU2_noalloc.

It must give mordred_filter_probe_group_by_GPU2's result word for word (and so
the oracle's), on two calls in a row on one GPU, the second with a smaller
input, with its compact probe tables in the fixed per-GPU scratch array that
the first call left behind.  Both calls compact at least one table, so the
second reads only what its own pre-pass wrote.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from oracle import queries as OQ

pytestmark = pytest.mark.gpu

SEG = oracle.SEGMENT_SIZE


def _run(fn, gq, n, res):
    import torch
    from mordred_amd import abi
    res.zero_()
    abi.call(fn, gq.cache.ptr, abi.ref(gq.fargs), C.pointer(gq.pargs), C.pointer(gq.gargs), n, res.data_ptr(),
             0, gq.fact_sg.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return res.cpu().numpy()


# (parts, fact rows): the first compacts the supplier table (1 bit) and the
# part table (16-bit brands + existence bitmap); the second, smaller, compacts
# only the supplier table
CALLS = [(300_000, 5 * SEG + 4_321), (100_000, 2 * SEG + 999)]


def test_noalloc_equals_pooled_entry_point_twice(cuda):
    import torch
    from mordred_amd import query as Q
    for i, (parts, n) in enumerate(CALLS):
        q = Q.q21(5_000, parts)
        data = Q.make_data(q, n, seed=90 + i, S=5_000, P=parts)
        assert n >= 16 * 5_000
        print(f"[gpu] no-alloc call {i}: {n} rows, part table {parts} slots", flush=True)
        cache = Q.load_cache(data, q, shuffle_seed=13 + i)
        gq = Q.GpuQuery(q, data, cache)
        gq.reset()
        gq.build()
        torch.cuda.synchronize()
        res = torch.zeros(q.total_val * 6, dtype=torch.int32, device="cuda")
        noalloc = _run("mordred_filter_probe_group_by_GPU2_noalloc", gq, n, res)
        pooled = _run("mordred_filter_probe_group_by_GPU2", gq, n, res)
        np.testing.assert_array_equal(noalloc, pooled)
        np.testing.assert_array_equal(noalloc, OQ.run_query(q, data, oracle.CONV_GPU))
        assert (oracle.res_sums(noalloc) != 0).sum() > 1000


def test_noalloc_refuses_tables_beyond_the_scratch(cuda):
    """A part hash table of 8,400,000 slots (100,000 part rows, so no key
    collides) probed by 129 x 2^20 fact positions -- a segment group that
    lists one cached segment 129 times -- gets a compact table of 16.8 MB,
    more than the 16 MiB array: the no-alloc entry point refuses it with
    hipErrorOutOfMemory before launching anything, and leaves the result
    untouched.  The pooled entry point runs the same call, and its answer is
    129 times the one-segment answer of the oracle."""
    import dataclasses

    import torch
    from mordred_amd import abi
    from mordred_amd import query as Q
    q = Q.q21(5_000, 100_000)
    q = dataclasses.replace(q, joins=[dataclasses.replace(j, dim_len=8_400_000) if j.slot == 3 else j
                                      for j in q.joins])
    data = Q.make_data(q, SEG, seed=95, S=5_000, P=100_000)
    cache = Q.load_cache(data, q, shuffle_seed=17)
    gq = Q.GpuQuery(q, data, cache)
    gq.reset()
    gq.build()
    reps = 129
    sg = torch.zeros(reps, dtype=torch.int16, device="cuda")
    n = reps * SEG
    assert n >= 16 * 8_400_000
    res = torch.zeros(q.total_val * 6, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    print("[gpu] no-alloc refusal: 8.4M-slot part table", flush=True)
    rc = abi.load().mordred_filter_probe_group_by_GPU2_noalloc(
        cache.ptr, abi.ref(gq.fargs), C.pointer(gq.pargs), C.pointer(gq.gargs), n, res.data_ptr(), 0,
        sg.data_ptr(), stream)
    torch.cuda.synchronize()
    assert rc == 2                      # hipErrorOutOfMemory
    assert not res.cpu().numpy().any()
    abi.call("mordred_filter_probe_group_by_GPU2", cache.ptr, abi.ref(gq.fargs), C.pointer(gq.pargs),
             C.pointer(gq.gargs), n, res.data_ptr(), 0, sg.data_ptr(), stream)
    torch.cuda.synchronize()
    got = res.cpu().numpy()
    one = OQ.run_query(q, data, oracle.CONV_GPU)
    np.testing.assert_array_equal(oracle.res_sums(got), reps * oracle.res_sums(one))
    np.testing.assert_array_equal(oracle.res_keys(got), oracle.res_keys(one))
