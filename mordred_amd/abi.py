"""ctypes binding of the C ABI declared in include/mordred_hip.h.

Struct layouts follow src/gpudb/KernelArgs.h (the device function-pointer
slots are kept as ignored ``void*``).  Every pointer argument is a DEVICE
pointer passed as an int (e.g. ``tensor.data_ptr()``) or ``None`` for NULL.
Entry points return a hipError_t code; :func:`check` turns a non-zero code
into :class:`MordredError`.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmordred_hip.so")
SEGMENT_SIZE = 1 << 20

VP = C.c_void_p
I = C.c_int


class MordredError(RuntimeError):
    """A mordred_* entry point returned a non-zero hipError_t."""

    def __init__(self, name: str, code: int):
        super().__init__(f"{name} failed with hipError_t {code}")
        self.code = code


class ProbeArgsGPU(C.Structure):
    """src/gpudb/KernelArgs.h:146-169."""
    _fields_ = ([(f"key_idx{i}", VP) for i in range(1, 5)] + [(f"ht{i}", VP) for i in range(1, 5)]
                + [(f"dim_len{i}", I) for i in range(1, 5)] + [(f"min_key{i}", I) for i in range(1, 5)])


class BuildArgsGPU(C.Structure):
    """src/gpudb/KernelArgs.h:196-205 (val_min carries the key minimum)."""
    _fields_ = [("key_idx", VP), ("val_idx", VP), ("num_slots", I), ("val_min", I), ("val_max", I)]


class GroupbyArgsGPU(C.Structure):
    """src/gpudb/KernelArgs.h:218-242."""
    _fields_ = ([("aggr_idx1", VP), ("aggr_idx2", VP)] + [(f"group_idx{i}", VP) for i in range(1, 5)]
                + [(f"min_val{i}", I) for i in range(1, 5)] + [(f"unique_val{i}", I) for i in range(1, 5)]
                + [("total_val", I), ("mode", I), ("d_group_func", VP)])


class FilterArgsGPU(C.Structure):
    """src/gpudb/KernelArgs.h:270-286."""
    _fields_ = [("filter_idx1", VP), ("filter_idx2", VP), ("compare1", I), ("compare2", I),
                ("compare3", I), ("compare4", I), ("mode1", I), ("mode2", I),
                ("d_filter_func1", VP), ("d_filter_func2", VP)]


class OffsetGPU(C.Structure):
    """src/gpudb/KernelArgs.h:306-315."""
    _fields_ = [("lo_off", VP)] + [(f"dim_off{i}", VP) for i in range(1, 5)]


_F, _P, _B, _G, _O = (C.POINTER(FilterArgsGPU), C.POINTER(ProbeArgsGPU), C.POINTER(BuildArgsGPU),
                      C.POINTER(GroupbyArgsGPU), C.POINTER(OffsetGPU))

# name -> argtypes (the trailing VP is the stream)
SIGNATURES = {
    "mordred_filter_probe_group_by_GPU2": [VP, _F, _P, _G, I, VP, I, VP, VP],
    "mordred_filter_probe_group_by_GPU2_noalloc": [VP, _F, _P, _G, I, VP, I, VP, VP],
    "mordred_filter_probe_group_by_GPU3": [VP, _O, _F, _P, _G, I, VP, VP],
    "mordred_probe_group_by_GPU2": [VP, _P, _G, I, VP, I, VP, VP],
    "mordred_probe_group_by_GPU3": [VP, _O, _P, _G, I, VP, VP],
    "mordred_groupByGPU": [VP, _O, _G, I, VP, VP],
    "mordred_filter_probe_GPU2": [VP, _F, _P, _O, I, VP, I, VP, VP],
    "mordred_filter_probe_GPU3": [VP, _O, _F, _P, _O, I, VP, VP],
    "mordred_probe_GPU2": [VP, _P, _O, I, VP, I, VP, VP],
    "mordred_probe_GPU3": [VP, _O, _P, _O, I, VP, VP],
    "mordred_filter_GPU2": [VP, _F, VP, I, VP, I, VP, VP],
    "mordred_filter_GPU3": [VP, VP, _F, VP, I, VP, VP],
    "mordred_build_GPU2": [VP, _F, _B, I, VP, I, VP, VP],
    "mordred_build_GPU3": [VP, VP, _F, _B, I, VP, VP],
    "mordred_build_GPU_minmax": [VP, _F, _B, I, VP, VP, VP, I, VP, VP],
    "mordred_build_GPU_minmax2": [VP, VP, _F, _B, I, VP, VP, VP, VP],
    "mordred_aggregationGPU": [VP, VP, _G, I, VP, VP],
    "mordred_probe_aggr_GPU2": [VP, _P, _G, I, VP, I, VP, VP],
    "mordred_probe_aggr_GPU3": [VP, _O, _P, _G, I, VP, VP],
    "mordred_filter_probe_aggr_GPU2": [VP, _F, _P, _G, I, VP, I, VP, VP],
    "mordred_filter_probe_aggr_GPU3": [VP, _O, _F, _P, _G, I, VP, VP],
    "mordred_build_GPU": [VP, VP, VP, I, I, I, I, VP, I, I, I, VP],
    "mordred_filter_probe_aggr_GPU": [VP, VP, I, I, I, I, I, I, VP, VP, I, I, VP, VP, I, VP, VP],
    "mordred_probe_group_by_GPU": ([VP] * 4 + [VP, I] * 4 + [I] * 4 + [VP, VP] + [I] * 4 + [I] * 4
                                   + [I, I, VP, VP]),
    "mordred_probe_table_register": [VP, I],
    "mordred_probe_table_unregister": [VP],
    "mordred_probe_table_invalidate": [VP],
}

_lib = None


def load() -> C.CDLL:
    """Load libmordred_hip.so; raise if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.mordred_version.argtypes = []
    lib.mordred_version.restype = C.c_char_p
    _lib = lib
    return lib


def check(name: str, code: int) -> None:
    if code != 0:
        raise MordredError(name, code)


def call(name: str, *args) -> None:
    """Call entry point `name` and raise MordredError on a non-zero return."""
    check(name, getattr(load(), name)(*args))


def ref(struct) -> "C._Pointer | None":
    """Pointer to a struct, or NULL for None."""
    return None if struct is None else C.pointer(struct)
