/*
 * mordred_engine.h -- host engine around the MI355X kernels (C ABI).
 *
 * The segment cache, the CPU/GPU split and the merge of the Mordred engine:
 *   - a column store in pinned host memory, cut into 2^20-row segments, with
 *     an HBM segment cache holding a subset of the segments
 *     (src/gpudb/CacheManager.cu:34-96, 306-361);
 *   - per query, fact segments are grouped by the residency of the query's
 *     columns (QueryOptimizer.cu:1674-1716); a group whose columns are all
 *     cached runs the fused GPU kernel, the others run on the engine's CPU
 *     operators (the role of src/gpudb/CPUProcessing.cu); the two partial
 *     results are merged (CPUProcessing.cu:1786-1819);
 *   - residency and placement follow the analytic cost model of
 *     src/gpudb/CostModel.cu:283-440 (bytes / bandwidth, GPU operators free).
 * On top of the reference's placement, MORDRED_PLACE_COST_SPLIT lets the cost
 * model send part of the non-resident segments over PCIe to the GPU (streamed
 * through staging slots of the cache, as the reference's on-demand path does,
 * QueryProcessing.cu:1001-1187) so that CPU work and transfers finish together.
 *
 * All entry points return 0 on success, a positive hipError_t from the HIP
 * runtime, or -1 for an invalid argument (with a message from
 * mordred_engine_last_error).  One engine is used by one host thread at a time.
 */
#ifndef MORDRED_ENGINE_H
#define MORDRED_ENGINE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mordred_engine mordred_engine;

/* One joined dimension (the per-table entries of QueryParams, KernelArgs.h:103-144). */
typedef struct mordred_join {
  int slot;                /* 1 supplier, 2 customer, 3 part, 4 date (table_id) */
  const char* fact_key;    /* fact foreign-key column */
  const char* dim_key;     /* dimension key column */
  int dim_len;             /* hash-table slots */
  int min_key;             /* HASH minimum */
  const char* filter_col;  /* dimension filter column or NULL */
  int filter_mode;         /* 0 EQ, 1 BETWEEN, 2 EQ-or-EQ */
  int c1, c2;
  const char* group_col;   /* value stored in the hash table, or NULL */
  int min_val, unique_val; /* dense group-index stride of this table */
} mordred_join;

typedef struct mordred_fact_filter {
  const char* col;         /* fact column, BETWEEN lo..hi */
  int lo, hi;
} mordred_fact_filter;

enum { MORDRED_KIND_AGGR = 0, MORDRED_KIND_GROUPBY = 1 };

typedef struct mordred_star_query {
  int njoins;
  mordred_join joins[4];
  int nfilters;
  mordred_fact_filter filters[2];
  const char* aggr1;
  const char* aggr2;       /* NULL allowed */
  int kind;                /* AGGR: SUM(a1*a2); GROUPBY: per-group SUM(a1-a2) */
  int total_val;           /* groups (1 for AGGR) */
} mordred_star_query;

enum {
  MORDRED_PLACE_RESIDENCY = 0,  /* reference: cached groups on GPU, partly cached ones mixed, the rest on CPU */
  MORDRED_PLACE_COST_SPLIT = 1, /* cost model splits non-resident segments CPU / PCIe+GPU */
  MORDRED_PLACE_GPU_ONLY = 2,   /* every non-resident segment streamed to the GPU */
  MORDRED_PLACE_CPU_ONLY = 3    /* every segment on the CPU operators */
};

/* Operator order of a partly cached segment group: which device runs each
 * filter, probe and the group-by, and in what order (each change of device
 * copies the row ids and dimension rows kept so far over PCIe, as
 * switch_device_fact, CPUGPUProcessing.cu:114-136).  A filter or probe runs
 * on the GPU iff its fact column is cached in every segment of the group. */
enum {
  /* every GPU filter and probe first, then one switch; the CPU operators
   * finish, group-by included (default) */
  MORDRED_ORDER_GPU_FIRST = 0,
  /* executeTableFact_v1 (QueryProcessing.cu:201-437): CPU filters, GPU
   * filters and probes, CPU probes, then the group-by on the GPU when its
   * fact columns are cached, else on the CPU */
  MORDRED_ORDER_CPU_FIRST = 1,
  /* executeTableFact_v2 (QueryProcessing.cu:676-914): GPU filters, CPU
   * filters and probes, GPU probes, then the group-by as above */
  MORDRED_ORDER_ALTERNATE = 2,
  /* main.cu:155-161 runs both fact pipelines and keeps the faster: the first
   * runs of a query with a given residency of its partly cached groups try
   * each order above once; later runs use the one whose run was fastest */
  MORDRED_ORDER_AUTO = 3
};
int mordred_engine_set_operator_order(mordred_engine* e, int order);

typedef struct mordred_run_stats {
  double total_ms;         /* whole query, builds to merged result */
  double gpu_ms;           /* GPU part: builds, cached groups, streamed groups */
  double cpu_ms;           /* CPU part: builds and CPU groups */
  double merge_ms;
  long long gpu_rows, cpu_rows, stream_rows;
  int gpu_segments, cpu_segments, stream_segments;
  double est_cpu_ms_per_segment;    /* analytic model */
  double est_stream_ms_per_segment; /* analytic model */
  int skipped_segments;             /* fact segments excluded by min/max */
  /* COST_SPLIT with a learned placement model (mordred_engine_set_learned_placement) */
  int plan_model;                   /* 0 analytic, 1 learned, 2 explored at random */
  int plan_streamed;                /* non-resident segments the plan streams */
  double plan_predicted_ms;         /* learned model's time for the plan (plan_model 1) */
  double plan_analytic_ms;          /* analytic model's time for the plan */
  /* partly cached segments run as mixed pipelines: the filters and probes whose
   * fact column is cached on the GPU (into offset lists), the rest of the
   * pipeline on the CPU operators over the survivors */
  int mixed_segments;
  long long mixed_rows;             /* fact rows of the mixed segments */
  long long mixed_survivors;        /* rows that left the GPU operators */
  double est_mixed_ms_per_segment;  /* analytic model (COST_SPLIT) */
  /* COST_SPLIT prefix streaming: of these non-resident segments only the
   * columns of the chosen fact filters and probes crossed PCIe; those
   * operators ran on the GPU, the CPU operators finished the survivors */
  int prefix_segments;
  long long prefix_rows, prefix_survivors;
  double est_prefix_ms_per_segment;      /* analytic model: transfers */
  double est_prefix_cpu_ms_per_segment;  /* analytic model: CPU operators */
  double suffix_ms;                  /* CPU operators over mixed / prefix survivors */
  int operator_order;                /* MORDRED_ORDER_* the mixed groups ran with */
  int order_switches;                /* device switches of ordered mixed pipelines (all chunks) */
  double cpu_build_ms;               /* the CPU hash-table builds (0 when none ran) */
  double est_cpu_build_ms;           /* their estimate in the split model (0 before any was measured) */
  /* bytes of the prefix columns that crossed PCIe bit-packed (values - column
   * minimum in the fewest bits, up to 24; unpacked on the device); a wider
   * column crosses as 4-byte values and adds nothing here */
  long long prefix_packed_bytes;
} mordred_run_stats;

/* Engine on HIP device `device` with an HBM segment cache of cache_segments
 * segments (4 MiB each) and `threads` CPU worker threads (0 = hardware). */
mordred_engine* mordred_engine_create(int device, int cache_segments, int threads);
void mordred_engine_destroy(mordred_engine* e);
const char* mordred_engine_last_error(const mordred_engine* e);
/* sizeof of a struct of this header by its name ("mordred_run_stats", ...),
 * or 0 for an unknown name: bindings check their copies of the layouts. */
long long mordred_engine_struct_size(const char* name);

/* Copy a column (n int32 values) into the engine's pinned host store.  With
 * gpu_resident set, every segment is also cached in HBM (dimension tables). */
int mordred_engine_add_column(mordred_engine* e, const char* name, const int* values, long long n,
                              int gpu_resident);
int mordred_engine_column_segments(const mordred_engine* e, const char* name);

/* Cache / evict one segment of a column (CacheManager.cu:306-361). */
int mordred_engine_cache_segment(mordred_engine* e, const char* name, int segment);
int mordred_engine_evict_segment(mordred_engine* e, const char* name, int segment);
/* 1 if the segment of the column is in HBM. */
int mordred_engine_is_cached(const mordred_engine* e, const char* name, int segment);

/* ---- on-disk columns, per-segment min/max, segment skipping -------------
 * Column files are headerless little-endian int32 arrays, one file per
 * column (common.h:137-201 loadColumn / storeColumn; the reference names them
 * <TABLE><index>, e.g. LINEORDER5, and LINEORDERSORT5 for the table sorted by
 * lo_orderdate, common.h:84-135).  Every column the engine holds carries a
 * [min, max] per 2^20-row segment, computed from the values as they are
 * copied or read (what minmax.cpp:28-45 computes offline).  A minmax file has
 * one "<min> <max>" line per segment (CacheManager.cu:154-184).
 * With skipping on (the default, as main.cu:21), mordred_engine_run and
 * mordred_engine_place leave out the fact segments that a query's predicates
 * exclude (QueryOptimizer.cu:1640-1662): the fact filters, and the
 * foreign-key range a dimension filter implies (sorted or clustered key
 * columns only).  Results do not change. */

/* Load n values (n = -1: the whole file) of a column file into the store, as
 * mordred_engine_add_column.  The file may hold more values; fewer is an error. */
int mordred_engine_load_column_file(mordred_engine* e, const char* name, const char* path, long long n,
                                    int gpu_resident);
int mordred_engine_store_column_file(mordred_engine* e, const char* name, const char* path);
/* Replace a column's skipping ranges with a minmax file's.  The file must have
 * one line per segment, min <= max, and cover the segment's values (a range
 * tighter than the data is refused, since skipping with it would drop rows). */
int mordred_engine_load_segment_minmax(mordred_engine* e, const char* name, const char* path);
int mordred_engine_store_segment_minmax(mordred_engine* e, const char* name, const char* path);
/* mins / maxs receive one entry per segment (mordred_engine_column_segments). */
int mordred_engine_segment_minmax(mordred_engine* e, const char* name, int* mins, int* maxs);
int mordred_engine_set_skipping(mordred_engine* e, int on);
/* Dimension table forms a shared-scan batch may choose (a mask of the
 * MORDRED_BATCH_FORM_* bits; default all).  Mask: per slot the 64-bit mask of
 * the queries whose filters the row passes; attribute: per slot the packed
 * attributes; dictionary: per slot a code of the row's attribute tuple,
 * decoded through a table in LDS.  A dimension takes the dictionary form when
 * allowed and its rows hold at most 1,024 distinct attribute tuples, else the
 * attribute form when allowed and applicable (tables above 1 MiB in mask
 * form, or any table without the MASK bit), else the mask form (always
 * available).  Every form gives the same results.  No counterpart in the
 * reference. */
enum {
  MORDRED_BATCH_FORM_MASK = 1,
  MORDRED_BATCH_FORM_ATTR = 2,
  MORDRED_BATCH_FORM_DICT = 4,
  MORDRED_BATCH_FORM_ALL = 7
};
int mordred_engine_set_batch_forms(mordred_engine* e, int forms);
/* Identical queries in one shared-scan batch or launch (default on): when two
 * active queries are the same query (every field equal) and the first gets
 * LDS accumulators, the second does not scan; the rows of the launch count
 * alike for both, so the first's sums, touched groups and keys are added into
 * the second's result rows as well at the end of the launch.  Each query
 * keeps its own result rows and coverage, so the answers are unchanged; a
 * request stream that repeats queries scans each repeated query once per
 * launch.  No counterpart in the reference. */
int mordred_engine_set_batch_sharing(mordred_engine* e, int on);
/* COST_SPLIT prefix streaming (default on): the GPU side of the split may
 * stream only the columns of a subset of the fact filters and probes of a
 * non-resident segment, run those operators, and hand the survivors to the
 * CPU operators, when the analytic model prices that below streaming the
 * whole segment (the subset it prices lowest). */
int mordred_engine_set_prefix_streaming(mordred_engine* e, int on);

/* ---- learned placement -------------------------------------------------
 * No counterpart in the reference, whose optimizer places segments with the
 * analytic model alone (CostModel.cu:283-440, QueryOptimizer.cu:1718-1747).
 * With a mode other than OFF, an fp32 MLP (mordred_costmodel.h) predicts a
 * run's time from query and plan features and ranks candidate plans:
 *   - MORDRED_PLACE_COST_SPLIT: stream k = n*c/8 of the n non-resident
 *     segments (c = 0..8, plus the analytic k);
 *   - the operator order of partly cached segments (GPU-first, CPU-first,
 *     alternate; the reference's executeTableFact_v1/v2, main.cu:155-161)
 *     when the engine's order is MORDRED_ORDER_AUTO, under any placement;
 *   - the segments mordred_engine_place keeps: a segment weighs what the
 *     model predicts caching one more segment saves the query, instead of
 *     the analytic CPU cost (the weights accumulate over calls, as the
 *     reference's CacheManager.cu:533-587 weights do).
 * It is trained on the measured times of the runs themselves, scored against
 * the analytic model on every new run, and chooses only while it is the more
 * accurate of the two (else the analytic plan, the engine's order rule and
 * the analytic weights apply).  EXPLORE also runs a random candidate one time
 * in four, so that the model sees more than the plans it picks; FORCE lets
 * the model choose as soon as it has been trained once, accurate or not.
 * Every plan, order and residency gives the same exact result. */
enum { MORDRED_LEARNED_OFF = 0, MORDRED_LEARNED_USE = 1, MORDRED_LEARNED_EXPLORE = 2, MORDRED_LEARNED_FORCE = 3 };

typedef struct mordred_placement_stats {
  long long samples;      /* runs observed */
  int scored;
  double learned_error;   /* EWMA |log error| on fresh runs */
  double analytic_error;
  int trusted;
  long long order_choices;    /* runs whose operator order the model changed */
  long long residency_calls;  /* mordred_engine_place calls weighted by the model */
} mordred_placement_stats;

int mordred_engine_set_learned_placement(mordred_engine* e, int mode);
int mordred_engine_placement_model_stats(mordred_engine* e, mordred_placement_stats* st);

/* Analytic-model bandwidths in bytes per second (CostModel.h:5-6 defaults
 * 42e9 and 12e9). */
int mordred_engine_set_bandwidth(mordred_engine* e, double bw_cpu, double bw_pci);

/* Cost-model residency: rank the fact segments of q by the cost the GPU saves
 * on them (CostModel::permute_cost weights) and cache the query's fact
 * columns of the top `fraction` of segments, evicting the rest. */
int mordred_engine_place(mordred_engine* e, const mordred_star_query* q, double fraction);

/* Run q.  res (total_val*6 ints, or 6 for AGGR) receives the merged result
 * (overwritten).  stats may be NULL. */
int mordred_engine_run(mordred_engine* e, const mordred_star_query* q, int placement, int* res,
                       mordred_run_stats* stats);

/* ---- shared-scan batches ------------------------------------------------
 * No counterpart in the reference, which runs one query at a time: up to 64
 * star queries execute in one persistent kernel that reads the fact columns
 * once.  Each dimension is probed once per fact row through a packed
 * attribute table: per hash slot a present bit and the attributes any query
 * of the batch filters or groups on (1, 2 or 4 bytes).  The queries' filters
 * on a dimension become per-attribute tables of 64-bit query masks (narrow
 * attributes) or per-query range compares; every query's result equals its
 * own mordred_engine_run (GPU convention) bit for bit.  The launch sequence
 * (clear, packed-table builds, scan) is captured in a hipGraph on first use
 * and replayed for every later batch of the same dataset and configuration.
 * Requirements: every fact column of the batch cached in HBM (all segments),
 * dimension columns GPU-resident, queries that join one dimension agree on
 * its fact key, key column, dim_len and min_key, dimension keys within
 * [min_key, min_key + dim_len) (no hash collisions), at most 3 filter and
 * group columns per dimension packing into 31 bits, at most 12 distinct fact
 * columns and 4 distinct aggregate expressions. */
#define MORDRED_BATCH_MAX_QUERIES 64

typedef struct mordred_batch_config {
  int block;      /* threads per workgroup: 256, 512 or 1024 */
  int vpt;        /* 4-row groups per thread: 1 or 2 (1024 threads: 1) */
  int lds_bytes;  /* dynamic LDS for descriptors + group arrays, <= 160 KiB */
} mordred_batch_config;

typedef struct mordred_batch_stats {
  double total_ms;           /* host: upload, graph launch, results back */
  double gpu_ms;             /* HIP events around the graph launch */
  int lds_queries;           /* queries whose group array sits in LDS */
  int global_queries;        /* queries using global atomics */
  int dead_queries;          /* empty predicate range: result stays zero */
  int graph_captured;        /* 1 if this call (re)captured the graph */
  mordred_batch_config config;
  double alg_bytes;          /* distinct fact columns read x 4 B x rows scanned */
  int lds32_queries;         /* of lds_queries: 32-bit LDS sums (a returning add per row) */
  int keys_in_loop_queries;  /* keys stored by the scan (a returning add per row) */
  int shared_queries;        /* identical to another query of the batch: took its sums */
} mordred_batch_stats;

typedef struct mordred_batch_tune_entry {
  mordred_batch_config config;
  double gpu_ms;             /* best of the timed runs, < 0 if the config failed */
} mordred_batch_tune_entry;

/* Run nq star queries as one batch.  res[i] receives query i's result
 * (total_val*6 ints, or 6 for AGGR), overwritten.  cfg NULL = the engine's
 * tuned configuration (or its default). */
int mordred_engine_batch(mordred_engine* e, const mordred_star_query* qs, int nq, int* const* res,
                         const mordred_batch_config* cfg, mordred_batch_stats* stats);

/* Autotune the batch launch on this batch: every candidate (threads per
 * workgroup x items per thread x LDS budget, which moves group arrays between
 * LDS and HBM) runs `reps` timed times after one warm-up; all candidates must
 * agree bit for bit.  The fastest becomes the engine's configuration and is
 * written to *best.  table (may be NULL) receives up to max_entries entries. */
int mordred_engine_batch_tune(mordred_engine* e, const mordred_star_query* qs, int nq, int reps,
                              mordred_batch_config* best, mordred_batch_tune_entry* table, int max_entries,
                              int* nentries);

/* ---- request loop -------------------------------------------------------
 * No counterpart in the reference.  Replays a stream of requests (query i
 * arrives arrival_ms[i] after the loop starts, ascending) in real time on
 * this thread and returns when every request is answered.
 *   MORDRED_SERVE_PLAIN: first come first served, each query alone through
 *     mordred_engine_run (its builds and its fused kernel).
 *   MORDRED_SERVE_CONTINUOUS: a circular shared scan.  The GPU walks the fact
 *     table in launches of consecutive segments; every launch runs the shared
 *     scan for all active queries.  A request is admitted at the next launch
 *     boundary when a query slot is free and its result rows (24 bytes per
 *     group) fit the HBM budget, and finishes once the scan has covered every
 *     segment exactly once for it (launches never cross an active query's
 *     starting segment), so its result equals its own run bit for bit.
 *     The loop plans the next launch (admission, plan, launch size) while the
 *     current one runs: the queries a launch completes are known when it is
 *     planned, so their slots go to the next plan, and their results are
 *     copied out behind the launch before the next one can reuse a slot.
 *     Launch sizes (segments per launch) are chosen per launch so that the
 *     predicted launch time stays within target_chunk_ms; predictions come
 *     from the learned cost model once it is trusted (mordred_costmodel.h),
 *     else from the analytic launch model.  The learned model is trained
 *     online on the measured launch times and persists in the engine.
 *     With lone_plain set, a request that arrives while no query is active
 *     and no other request is waiting runs alone through mordred_engine_run
 *     (its builds and its fused kernel, the same answer): no other query
 *     could share its scan, and one query's fused kernel finishes sooner than
 *     a pass of the shared scan.  Requests arriving meanwhile wait for it.
 * The fact columns must be cached in HBM (all segments) for both modes. */
enum { MORDRED_SERVE_PLAIN = 0, MORDRED_SERVE_CONTINUOUS = 1 };

typedef struct mordred_serve_opts {
  int mode;
  int max_active;             /* query slots (<= 64) */
  int chunk_segments;         /* > 0: fixed launch size; 0: chosen per launch */
  double target_chunk_ms;     /* launch-time target for the chosen launch size */
  double hbm_budget_bytes;    /* admission budget for per-query device state */
  int learned;                /* 1: use the learned cost model when trusted */
  mordred_batch_config config;  /* block 0: the engine's batch configuration */
  int lone_plain;             /* continuous: 1 = a lone request on an idle loop runs alone */
} mordred_serve_opts;

typedef struct mordred_serve_times {
  double arrival_ms, admit_ms, done_ms;  /* since the loop started */
  int launches;                          /* launches the request took part in */
} mordred_serve_times;

typedef struct mordred_serve_stats {
  double wall_ms;
  long long launches;
  double gpu_ms;                  /* sum of launch times (events) */
  int max_active_seen;
  long long learned_choices;      /* launch sizes chosen with the learned model */
  long long analytic_choices;
  double learned_error;           /* EWMA |log error| of each model on launches */
  double analytic_error;
  int learned_trusted;
  double peak_state_bytes;        /* admitted per-query device state, at most */
  double alg_bytes;               /* continuous: sum over launches of the distinct
                                     fact columns read x 4 B x rows scanned */
  long long lone_requests;        /* continuous: requests run alone (lone_plain); launches,
                                     gpu_ms and alg_bytes count the shared-scan launches only */
  long long shared_queries;       /* continuous: sum over launches of the active queries that
                                     took an identical query's sums (mordred_engine_set_batch_sharing) */
} mordred_serve_stats;

int mordred_engine_serve(mordred_engine* e, const mordred_star_query* qs, int n, const double* arrival_ms,
                         int* const* res, const mordred_serve_opts* opts, mordred_serve_times* times,
                         mordred_serve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* MORDRED_ENGINE_H */
