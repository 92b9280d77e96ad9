"""ctypes binding of the host engine (include/mordred_engine.h).

:class:`Engine` owns a column store (pinned host memory), an HBM segment cache
and the CPU worker pool; :func:`star_query` converts a :class:`query.StarQuery`
into the C struct.  ``Engine(device=-1)`` is a host-only engine that runs the
CPU operators without any HIP call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

PLACE_RESIDENCY, PLACE_COST_SPLIT, PLACE_GPU_ONLY, PLACE_CPU_ONLY = 0, 1, 2, 3
# operator order of partly cached segment groups
ORDER_GPU_FIRST, ORDER_CPU_FIRST, ORDER_ALTERNATE, ORDER_AUTO = 0, 1, 2, 3
KIND_AGGR, KIND_GROUPBY = 0, 1


class Join(C.Structure):
    _fields_ = [("slot", C.c_int), ("fact_key", C.c_char_p), ("dim_key", C.c_char_p), ("dim_len", C.c_int),
                ("min_key", C.c_int), ("filter_col", C.c_char_p), ("filter_mode", C.c_int), ("c1", C.c_int),
                ("c2", C.c_int), ("group_col", C.c_char_p), ("min_val", C.c_int), ("unique_val", C.c_int)]


class FactFilter(C.Structure):
    _fields_ = [("col", C.c_char_p), ("lo", C.c_int), ("hi", C.c_int)]


class StarQueryC(C.Structure):
    _fields_ = [("njoins", C.c_int), ("joins", Join * 4), ("nfilters", C.c_int), ("filters", FactFilter * 2),
                ("aggr1", C.c_char_p), ("aggr2", C.c_char_p), ("kind", C.c_int), ("total_val", C.c_int)]


class RunStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("gpu_ms", C.c_double), ("cpu_ms", C.c_double), ("merge_ms", C.c_double),
                ("gpu_rows", C.c_longlong), ("cpu_rows", C.c_longlong), ("stream_rows", C.c_longlong),
                ("gpu_segments", C.c_int), ("cpu_segments", C.c_int), ("stream_segments", C.c_int),
                ("est_cpu_ms_per_segment", C.c_double), ("est_stream_ms_per_segment", C.c_double),
                ("skipped_segments", C.c_int), ("plan_model", C.c_int), ("plan_streamed", C.c_int),
                ("plan_predicted_ms", C.c_double), ("plan_analytic_ms", C.c_double),
                ("mixed_segments", C.c_int), ("mixed_rows", C.c_longlong), ("mixed_survivors", C.c_longlong),
                ("est_mixed_ms_per_segment", C.c_double), ("prefix_segments", C.c_int),
                ("prefix_rows", C.c_longlong), ("prefix_survivors", C.c_longlong),
                ("est_prefix_ms_per_segment", C.c_double), ("est_prefix_cpu_ms_per_segment", C.c_double),
                ("suffix_ms", C.c_double), ("operator_order", C.c_int), ("order_switches", C.c_int),
                ("cpu_build_ms", C.c_double), ("est_cpu_build_ms", C.c_double),
                ("prefix_packed_bytes", C.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PlacementStats(C.Structure):
    _fields_ = [("samples", C.c_longlong), ("scored", C.c_int), ("learned_error", C.c_double),
                ("analytic_error", C.c_double), ("trusted", C.c_int), ("order_choices", C.c_longlong),
                ("residency_calls", C.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


LEARNED_OFF, LEARNED_USE, LEARNED_EXPLORE, LEARNED_FORCE = 0, 1, 2, 3


class BatchConfig(C.Structure):
    _fields_ = [("block", C.c_int), ("vpt", C.c_int), ("lds_bytes", C.c_int)]

    def as_tuple(self):
        return (self.block, self.vpt, self.lds_bytes)


class BatchStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("gpu_ms", C.c_double), ("lds_queries", C.c_int),
                ("global_queries", C.c_int), ("dead_queries", C.c_int), ("graph_captured", C.c_int),
                ("config", BatchConfig), ("alg_bytes", C.c_double), ("lds32_queries", C.c_int),
                ("keys_in_loop_queries", C.c_int), ("shared_queries", C.c_int)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "config"}
        d["config"] = self.config.as_tuple()
        return d


class BatchTuneEntry(C.Structure):
    _fields_ = [("config", BatchConfig), ("gpu_ms", C.c_double)]


SERVE_PLAIN, SERVE_CONTINUOUS = 0, 1


class ServeOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("max_active", C.c_int), ("chunk_segments", C.c_int),
                ("target_chunk_ms", C.c_double), ("hbm_budget_bytes", C.c_double), ("learned", C.c_int),
                ("config", BatchConfig), ("lone_plain", C.c_int)]


class ServeTimes(C.Structure):
    _fields_ = [("arrival_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double),
                ("launches", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("wall_ms", C.c_double), ("launches", C.c_longlong), ("gpu_ms", C.c_double),
                ("max_active_seen", C.c_int), ("learned_choices", C.c_longlong), ("analytic_choices", C.c_longlong),
                ("learned_error", C.c_double), ("analytic_error", C.c_double), ("learned_trusted", C.c_int),
                ("peak_state_bytes", C.c_double), ("alg_bytes", C.c_double), ("lone_requests", C.c_longlong),
                ("shared_queries", C.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def _b(s):
    return None if s is None else s.encode()


def star_query(q) -> StarQueryC:
    """C struct of a query.StarQuery (strings are encoded; keep the result alive)."""
    c = StarQueryC()
    c.njoins = len(q.joins)
    for i, j in enumerate(q.joins):
        c.joins[i] = Join(j.slot, _b(j.fact_key), _b(j.dim_key), j.dim_len, j.min_key, _b(j.filter_col),
                          j.filter_mode, j.c1, j.c2, _b(j.group_col), j.min_val, j.unique_val)
    c.nfilters = len(q.fact_filters)
    for i, (col, lo, hi) in enumerate(q.fact_filters):
        c.filters[i] = FactFilter(_b(col), lo, hi)
    c.aggr1, c.aggr2 = _b(q.aggr[0]), _b(q.aggr[1])
    c.kind = KIND_AGGR if q.kind == "aggr" else KIND_GROUPBY
    c.total_val = max(1, q.total_val)
    return c


_SIGS = {
    "mordred_engine_create": ([C.c_int, C.c_int, C.c_int], C.c_void_p),
    "mordred_engine_destroy": ([C.c_void_p], None),
    "mordred_engine_last_error": ([C.c_void_p], C.c_char_p),
    "mordred_engine_add_column": ([C.c_void_p, C.c_char_p, C.c_void_p, C.c_longlong, C.c_int], C.c_int),
    "mordred_engine_column_segments": ([C.c_void_p, C.c_char_p], C.c_int),
    "mordred_engine_cache_segment": ([C.c_void_p, C.c_char_p, C.c_int], C.c_int),
    "mordred_engine_evict_segment": ([C.c_void_p, C.c_char_p, C.c_int], C.c_int),
    "mordred_engine_is_cached": ([C.c_void_p, C.c_char_p, C.c_int], C.c_int),
    "mordred_engine_set_bandwidth": ([C.c_void_p, C.c_double, C.c_double], C.c_int),
    "mordred_engine_load_column_file": ([C.c_void_p, C.c_char_p, C.c_char_p, C.c_longlong, C.c_int], C.c_int),
    "mordred_engine_store_column_file": ([C.c_void_p, C.c_char_p, C.c_char_p], C.c_int),
    "mordred_engine_load_segment_minmax": ([C.c_void_p, C.c_char_p, C.c_char_p], C.c_int),
    "mordred_engine_store_segment_minmax": ([C.c_void_p, C.c_char_p, C.c_char_p], C.c_int),
    "mordred_engine_segment_minmax": ([C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p], C.c_int),
    "mordred_engine_set_skipping": ([C.c_void_p, C.c_int], C.c_int),
    "mordred_engine_set_batch_forms": ([C.c_void_p, C.c_int], C.c_int),
    "mordred_engine_set_batch_sharing": ([C.c_void_p, C.c_int], C.c_int),
    "mordred_engine_set_prefix_streaming": ([C.c_void_p, C.c_int], C.c_int),
    "mordred_engine_set_operator_order": ([C.c_void_p, C.c_int], C.c_int),
    "mordred_engine_struct_size": ([C.c_char_p], C.c_longlong),
    "mordred_engine_set_learned_placement": ([C.c_void_p, C.c_int], C.c_int),
    "mordred_engine_placement_model_stats": ([C.c_void_p, C.POINTER(PlacementStats)], C.c_int),
    "mordred_engine_place": ([C.c_void_p, C.POINTER(StarQueryC), C.c_double], C.c_int),
    "mordred_engine_run": ([C.c_void_p, C.POINTER(StarQueryC), C.c_int, C.c_void_p, C.POINTER(RunStats)], C.c_int),
    "mordred_engine_batch": ([C.c_void_p, C.POINTER(StarQueryC), C.c_int, C.POINTER(C.c_void_p),
                              C.POINTER(BatchConfig), C.POINTER(BatchStats)], C.c_int),
    "mordred_engine_batch_tune": ([C.c_void_p, C.POINTER(StarQueryC), C.c_int, C.c_int, C.POINTER(BatchConfig),
                                   C.POINTER(BatchTuneEntry), C.c_int, C.POINTER(C.c_int)], C.c_int),
    "mordred_engine_serve": ([C.c_void_p, C.POINTER(StarQueryC), C.c_int, C.POINTER(C.c_double),
                              C.POINTER(C.c_void_p), C.POINTER(ServeOpts), C.POINTER(ServeTimes),
                              C.POINTER(ServeStats)], C.c_int),
}


def _lib():
    lib = abi.load()
    if not getattr(lib, "_engine_bound", False):
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        lib._engine_bound = True
    return lib


class EngineError(RuntimeError):
    pass


class Engine:
    def __init__(self, device: int = 0, cache_segments: int = 0, threads: int = 0):
        self.lib = _lib()
        self.h = self.lib.mordred_engine_create(device, cache_segments, threads)
        if not self.h:
            raise EngineError(f"mordred_engine_create(device={device}, cache_segments={cache_segments}) failed")
        self.device = device

    def close(self):
        if self.h:
            self.lib.mordred_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise EngineError(f"{what}: rc={rc} {self.lib.mordred_engine_last_error(self.h).decode()}")

    def add_column(self, name: str, values: np.ndarray, gpu_resident: bool = False):
        v = np.ascontiguousarray(values, dtype=np.int32)
        self._check(self.lib.mordred_engine_add_column(self.h, name.encode(), v.ctypes.data, len(v),
                                                       int(gpu_resident)), f"add_column {name}")

    def add_data(self, data, query):
        """Every column a StarQuery reads: fact columns in the host store,
        dimension columns cached in HBM (all of them) unless host-only."""
        for c in query.fact_columns():
            self.add_column(c, data.fact[c])
        seen = set()
        for j in query.joins:
            for c in (j.dim_key, j.filter_col, j.group_col):
                if c and c not in seen:
                    seen.add(c)
                    self.add_column(c, data.tables[j.table][c], gpu_resident=self.device >= 0)

    def add_mix_data(self, data, queries, cache_fact: bool = True):
        """Every column a list of StarQuery reads: fact columns (all segments
        cached in HBM with cache_fact) and GPU-resident dimension columns."""
        seen = set()
        for q in queries:
            for c in q.fact_columns():
                if c not in seen:
                    seen.add(c)
                    self.add_column(c, data.fact[c], gpu_resident=cache_fact)
        for q in queries:
            for j in q.joins:
                for c in (j.dim_key, j.filter_col, j.group_col):
                    if c and c not in seen:
                        seen.add(c)
                        self.add_column(c, data.tables[j.table][c], gpu_resident=self.device >= 0)

    def segments(self, name: str) -> int:
        return self.lib.mordred_engine_column_segments(self.h, name.encode())

    def cache_segment(self, name: str, seg: int):
        self._check(self.lib.mordred_engine_cache_segment(self.h, name.encode(), seg), "cache_segment")

    def evict_segment(self, name: str, seg: int):
        self._check(self.lib.mordred_engine_evict_segment(self.h, name.encode(), seg), "evict_segment")

    def is_cached(self, name: str, seg: int) -> bool:
        return self.lib.mordred_engine_is_cached(self.h, name.encode(), seg) == 1

    def set_bandwidth(self, bw_cpu: float, bw_pci: float):
        self._check(self.lib.mordred_engine_set_bandwidth(self.h, bw_cpu, bw_pci), "set_bandwidth")

    # ---- on-disk columns, per-segment min/max, skipping
    def load_column_file(self, name: str, path, n: int = -1, gpu_resident: bool = False):
        """Headerless little-endian int32 column file (n = -1: the whole file)."""
        self._check(self.lib.mordred_engine_load_column_file(self.h, name.encode(), str(path).encode(), n,
                                                             int(gpu_resident)), f"load_column_file {name}")

    def store_column_file(self, name: str, path):
        self._check(self.lib.mordred_engine_store_column_file(self.h, name.encode(), str(path).encode()),
                    f"store_column_file {name}")

    def load_segment_minmax(self, name: str, path):
        self._check(self.lib.mordred_engine_load_segment_minmax(self.h, name.encode(), str(path).encode()),
                    f"load_segment_minmax {name}")

    def store_segment_minmax(self, name: str, path):
        self._check(self.lib.mordred_engine_store_segment_minmax(self.h, name.encode(), str(path).encode()),
                    f"store_segment_minmax {name}")

    def segment_minmax(self, name: str):
        """(mins, maxs) per segment, the ranges skipping uses."""
        n = self.segments(name)
        if n < 0:
            raise EngineError(f"unknown column {name}")
        mins = np.zeros(max(1, n), dtype=np.int32)
        maxs = np.zeros(max(1, n), dtype=np.int32)
        self._check(self.lib.mordred_engine_segment_minmax(self.h, name.encode(), mins.ctypes.data, maxs.ctypes.data),
                    f"segment_minmax {name}")
        return mins[:n], maxs[:n]

    def set_skipping(self, on: bool):
        self._check(self.lib.mordred_engine_set_skipping(self.h, int(on)), "set_skipping")

    # dimension table forms a batch may use (include/mordred_engine.h)
    FORM_MASK, FORM_ATTR, FORM_DICT, FORM_ALL = 1, 2, 4, 7

    def set_batch_forms(self, forms: int):
        self._check(self.lib.mordred_engine_set_batch_forms(self.h, int(forms)), "set_batch_forms")

    def set_batch_sharing(self, on: bool):
        self._check(self.lib.mordred_engine_set_batch_sharing(self.h, int(bool(on))), "set_batch_sharing")

    def set_prefix_streaming(self, on: bool):
        self._check(self.lib.mordred_engine_set_prefix_streaming(self.h, int(on)), "set_prefix_streaming")

    def set_operator_order(self, order: int):
        """ORDER_GPU_FIRST (default), ORDER_CPU_FIRST (the reference's
        executeTableFact_v1), ORDER_ALTERNATE (executeTableFact_v2), or
        ORDER_AUTO (try each once per query and residency, keep the fastest)."""
        self._check(self.lib.mordred_engine_set_operator_order(self.h, int(order)), "set_operator_order")

    # ---- learned placement
    def set_learned_placement(self, mode: int):
        """LEARNED_OFF, LEARNED_USE (plans, operator orders under ORDER_AUTO and
        place()'s segment weights from the MLP while it is the more accurate
        model), LEARNED_EXPLORE (also a random plan one run in four) or
        LEARNED_FORCE (the MLP chooses once trained, accurate or not)."""
        self._check(self.lib.mordred_engine_set_learned_placement(self.h, mode), "set_learned_placement")

    def placement_model_stats(self) -> dict:
        st = PlacementStats()
        self._check(self.lib.mordred_engine_placement_model_stats(self.h, C.byref(st)), "placement_model_stats")
        return st.as_dict()

    def place(self, query, fraction: float):
        qc = star_query(query)
        self._check(self.lib.mordred_engine_place(self.h, C.byref(qc), fraction), "place")

    def run(self, query, placement: int = PLACE_RESIDENCY):
        qc = star_query(query)
        res = np.zeros(max(1, query.total_val) * 6, dtype=np.int32)
        st = RunStats()
        self._check(self.lib.mordred_engine_run(self.h, C.byref(qc), placement, res.ctypes.data, C.byref(st)), "run")
        return res, st.as_dict()

    # ---- shared-scan batches
    def _queries(self, queries):
        arr = (StarQueryC * len(queries))()
        for i, q in enumerate(queries):
            arr[i] = star_query(q)
        return arr

    def batch(self, queries, config=None):
        """Results (one int32 array per query) and stats of one batch."""
        arr = self._queries(queries)
        outs = [np.zeros(max(1, q.total_val) * 6, dtype=np.int32) for q in queries]
        ptrs = (C.c_void_p * len(queries))(*[o.ctypes.data for o in outs])
        st = BatchStats()
        cfg = None if config is None else C.byref(BatchConfig(*config))
        self._check(self.lib.mordred_engine_batch(self.h, arr, len(queries), ptrs, cfg, C.byref(st)), "batch")
        return outs, st.as_dict()

    def batch_tune(self, queries, reps: int = 2):
        """(best config tuple, [(config tuple, gpu_ms)]) of the batch autotuner."""
        arr = self._queries(queries)
        best = BatchConfig()
        table = (BatchTuneEntry * 64)()
        n = C.c_int(0)
        self._check(self.lib.mordred_engine_batch_tune(self.h, arr, len(queries), reps, C.byref(best), table, 64,
                                                       C.byref(n)), "batch_tune")
        return best.as_tuple(), [(table[i].config.as_tuple(), table[i].gpu_ms) for i in range(n.value)]

    # ---- request loop
    def serve(self, queries, arrivals_ms, mode=SERVE_CONTINUOUS, max_active: int = 64, chunk_segments: int = 0,
              target_chunk_ms: float = 8.0, hbm_budget_bytes: float = 200e9, learned: bool = True, config=None,
              lone_plain: bool = False):
        """Replay requests (query i arriving arrivals_ms[i] after the start) in
        real time.  Returns (results, per-request times, stats)."""
        n = len(queries)
        arr = self._queries(queries) if n else (StarQueryC * 1)()
        outs = [np.zeros(max(1, q.total_val) * 6, dtype=np.int32) for q in queries]
        ptrs = (C.c_void_p * max(1, n))(*[o.ctypes.data for o in outs])
        arrivals = (C.c_double * max(1, n))(*[float(a) for a in arrivals_ms])
        o = ServeOpts(mode, max_active, chunk_segments, target_chunk_ms, hbm_budget_bytes, int(learned),
                      BatchConfig(*(config or (0, 0, 0))), int(lone_plain))
        times = (ServeTimes * max(1, n))()
        st = ServeStats()
        self._check(self.lib.mordred_engine_serve(self.h, arr, n, arrivals, ptrs, C.byref(o), times, C.byref(st)),
                    "serve")
        t = [{"arrival_ms": times[i].arrival_ms, "admit_ms": times[i].admit_ms, "done_ms": times[i].done_ms,
              "launches": times[i].launches} for i in range(n)]
        return outs, t, st.as_dict()
