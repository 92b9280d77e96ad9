"""ORACLE -- test infrastructure, never the product.

ctypes binding of ``oracle/liboracle.so``, the plain-C restatement of the
reference engine's CPU path (src/gpudb/CPUProcessing.cu).  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package; the product (``mordred_amd``) never does.

Arrays are numpy int32 (hash tables, result buffers, offsets) or int16
(segment groups).  ``None`` stands for a NULL column.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
SEGMENT_SIZE = 1 << 20
CONV_CPU = 0
CONV_GPU = 1

_IP = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    """Compile liboracle.so from cpu_path.c with gcc + OpenMP."""
    src = os.path.join(HERE, "cpu_path.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB_PATH


class FilterArgs(C.Structure):
    """src/gpudb/KernelArgs.h:288-304 (filterArgsCPU)."""
    _fields_ = [("filter_col1", _IP), ("filter_col2", _IP),
                ("compare1", C.c_int), ("compare2", C.c_int), ("compare3", C.c_int), ("compare4", C.c_int),
                ("mode1", C.c_int), ("mode2", C.c_int),
                ("h_filter_func1", C.c_void_p), ("h_filter_func2", C.c_void_p)]


class ProbeArgs(C.Structure):
    """src/gpudb/KernelArgs.h:171-194 (probeArgsCPU)."""
    _fields_ = ([(f"key_col{i}", _IP) for i in range(1, 5)] + [(f"ht{i}", _IP) for i in range(1, 5)]
                + [(f"dim_len{i}", C.c_int) for i in range(1, 5)] + [(f"min_key{i}", C.c_int) for i in range(1, 5)])


class BuildArgs(C.Structure):
    """src/gpudb/KernelArgs.h:207-216 (buildArgsCPU)."""
    _fields_ = [("key_col", _IP), ("val_col", _IP), ("num_slots", C.c_int), ("val_min", C.c_int),
                ("val_max", C.c_int)]


class GroupbyArgs(C.Structure):
    """src/gpudb/KernelArgs.h:244-268 (groupbyArgsCPU)."""
    _fields_ = ([("aggr_col1", _IP), ("aggr_col2", _IP)] + [(f"group_col{i}", _IP) for i in range(1, 5)]
                + [(f"min_val{i}", C.c_int) for i in range(1, 5)] + [(f"unique_val{i}", C.c_int) for i in range(1, 5)]
                + [("total_val", C.c_int), ("mode", C.c_int), ("h_group_func", C.c_void_p)])


class Offset(C.Structure):
    """src/gpudb/KernelArgs.h:317-326 (offsetCPU)."""
    _fields_ = [("h_lo_off", _IP)] + [(f"h_dim_off{i}", _IP) for i in range(1, 5)]


def iptr(a):
    """int* of a C-contiguous int32 numpy array, or NULL for None."""
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_IP)


def sptr(a):
    if a is None:
        return None
    assert a.dtype == np.int16 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_short))


def _load():
    lib = C.CDLL(build())
    S, I, P, SP = C.Structure, C.c_int, _IP, C.POINTER(C.c_short)
    F, Pr, B, G, O = FilterArgs, ProbeArgs, BuildArgs, GroupbyArgs, Offset
    sig = {
        "oracle_filter_probe_CPU": [F, Pr, O, I, P, I, SP, I],
        "oracle_filter_probe_CPU2": [O, F, Pr, O, I, P, I, I],
        "oracle_probe_CPU": [Pr, O, I, P, I, SP, I],
        "oracle_probe_CPU2": [O, Pr, O, I, P, I, I],
        "oracle_probe_group_by_CPU": [Pr, G, I, P, I, SP, I],
        "oracle_probe_group_by_CPU2": [O, Pr, G, I, P, I, I],
        "oracle_filter_probe_group_by_CPU": [F, Pr, G, I, P, I, SP, I],
        "oracle_filter_probe_group_by_CPU2": [O, F, Pr, G, I, P, I, I],
        "oracle_build_CPU": [F, B, I, P, I, SP, I],
        "oracle_build_CPU2": [P, F, B, I, P, I, I],
        "oracle_filter_CPU": [F, P, I, P, I, SP, I],
        "oracle_filter_CPU2": [P, F, P, I, P, I, I],
        "oracle_groupByCPU": [O, G, I, P, I],
        "oracle_aggregationCPU": [P, G, I, P],
        "oracle_probe_aggr_CPU": [Pr, G, I, P, I, SP],
        "oracle_probe_aggr_CPU2": [O, Pr, G, I, P, I],
        "oracle_filter_probe_aggr_CPU": [F, Pr, G, I, P, I, SP, I],
        "oracle_filter_probe_aggr_CPU2": [O, F, Pr, G, I, P, I, I],
        "oracle_merge": [P, P, I],
        "oracle_build_CPU_minmax": [F, B, I, P, P, P, I, SP, I],
        "oracle_build_CPU_minmax2": [P, F, B, I, P, P, P, I, I],
        "oracle_set_threads": [I],
        "oracle_get_threads": [],
    }
    del S
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int if name == "oracle_get_threads" else None
    return lib


lib = _load()


def set_threads(n: int) -> None:
    lib.oracle_set_threads(int(n))


def get_threads() -> int:
    return int(lib.oracle_get_threads())


def res_sums(res: np.ndarray) -> np.ndarray:
    """int64 sums of a 6-int-per-group result buffer (ints 4-5 of each group)."""
    return res.reshape(-1, 6)[:, 4:6].copy().view(np.int64).reshape(-1)


def res_keys(res: np.ndarray) -> np.ndarray:
    return res.reshape(-1, 6)[:, 0:4].copy()


def res_scalar(res: np.ndarray) -> int:
    """int64 scalar aggregate stored at ints 4-5 (src/gpudb/GPUProcessing.h:1774)."""
    return int(res[4:6].copy().view(np.int64)[0])
