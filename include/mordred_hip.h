/*
 * mordred_hip.h -- C ABI of the MI355X (gfx950) query-execution kernels.
 *
 * Drop-in replacement for the kernel layer of the Mordred engine: the
 * __global__ templates of src/gpudb/GPUProcessing.h, launched today from
 * src/gpudb/CPUGPUProcessing.cu as  kernel<128,4><<<ceil(n/512),128,0,stream>>>.
 * Each entry point below replaces one such launch (reference kernel cited at
 * each declaration); the launch shape is internal to this library.
 *
 * Argument structs keep the field order of src/gpudb/KernelArgs.h so that the
 * engine's argument assembly (CPUGPUProcessing.cu:233-297) is unchanged.  The
 * device-function-pointer fields of the reference (filter / group functions
 * fetched with cudaMemcpyFromSymbol) are kept as ignored pointer-sized slots:
 * the reference kernels on this path never call them (the fused kernels
 * hard-code BETWEEN filters, a1*a2 and a1-a2), and the filter/build kernels
 * select the predicate with the mode fields (0 EQ, 1 BETWEEN, 2 EQ-or-EQ).
 *
 * Conventions shared by every entry point (same as the reference):
 *  - `gpuCache` is the segment cache: cache slot s holds one column segment
 *    of MORDRED_SEGMENT_SIZE int32 values at gpuCache + s*MORDRED_SEGMENT_SIZE.
 *    Every *_idx pointer is a DEVICE segment table: idx[segment] = cache slot
 *    (CacheManager.cu:246-257).  Slot arithmetic is 64-bit here, so the cache
 *    may exceed the reference's 2^31-int limit.
 *  - Segment mode (*_GPU2): position i < num_tuples of the segment group is
 *    the row  segment_group[i >> 20] * 2^20 + (i & (2^20-1)); with a NULL
 *    segment_group the segment is (start_offset + i) >> 20 and start_offset
 *    must be a multiple of 2^20.  `segment_group` is a DEVICE array of short.
 *  - Offset mode (*_GPU3): rows come from a DEVICE list of int32 row ids.
 *  - Hash tables are 2*num_slots int32, caller-zeroed, slot h =
 *    [value, row+1] (BlockLibrary.cuh:613-614); HASH(k,len,min) = (k-min)%len.
 *  - Results accumulate into caller-zeroed buffers: group-by results are
 *    total_val groups of 6 int32 (4 keys, then an int64 sum in ints 4-5);
 *    scalar aggregates are the int64 at ints 4-5.  Compaction output space is
 *    reserved with atomics on the caller-zeroed DEVICE counter `total`.
 *  - All launches are asynchronous on `stream` (NULL = default stream).  The
 *    return value is a hipError_t code: 0 on success, hipErrorInvalidValue (1)
 *    for arguments the kernels cannot honour (checked before any launch),
 *    otherwise the launch error.  Nothing here calls exit().
 *  - Entry points are re-entrant (callable from several threads and streams).
 *    Process-wide state, all of it internal and thread-safe: per device, a
 *    stream-ordered memory pool from which a fused call takes the scratch of
 *    its compact probe tables (allocated and freed in the call's stream
 *    order; the pool keeps freed blocks for reuse); a cache of occupancy
 *    queries for launch shapes; the fixed scratch array of the no-alloc entry
 *    point; and the probe tables a caller registers (declared below).  A
 *    caller that must own every allocation uses the no-alloc entry point or
 *    registers its hash tables.
 */
#ifndef MORDRED_HIP_H
#define MORDRED_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MORDRED_SEGMENT_SIZE 1048576 /* src/gpudb/common.h:73 */

/* Same type as hipStream_t (hip/hip_runtime_api.h). */
typedef struct ihipStream_t* mordred_stream_t;

/* src/gpudb/KernelArgs.h:146-169 */
typedef struct probeArgsGPU {
  int* key_idx1; int* key_idx2; int* key_idx3; int* key_idx4;
  int* ht1; int* ht2; int* ht3; int* ht4;
  int dim_len1; int dim_len2; int dim_len3; int dim_len4;
  int min_key1; int min_key2; int min_key3; int min_key4;
} probeArgsGPU;

/* src/gpudb/KernelArgs.h:196-205.  Callers pass {key, val, dim_len, key min}
 * (CPUGPUProcessing.cu:1606-1609): val_min is the KEY minimum. */
typedef struct buildArgsGPU {
  int* key_idx; int* val_idx;
  int num_slots; int val_min; int val_max;
} buildArgsGPU;

/* src/gpudb/KernelArgs.h:218-242 */
typedef struct groupbyArgsGPU {
  int* aggr_idx1; int* aggr_idx2;
  int* group_idx1; int* group_idx2; int* group_idx3; int* group_idx4;
  int min_val1; int min_val2; int min_val3; int min_val4;
  int unique_val1; int unique_val2; int unique_val3; int unique_val4;
  int total_val; int mode;
  void* d_group_func; /* ignored: group-by is a1-a2, aggregation a1*a2 */
} groupbyArgsGPU;

/* src/gpudb/KernelArgs.h:270-286 */
typedef struct filterArgsGPU {
  int* filter_idx1; int* filter_idx2;
  int compare1; int compare2; int compare3; int compare4;
  int mode1; int mode2;
  void* d_filter_func1; /* ignored */
  void* d_filter_func2; /* ignored */
} filterArgsGPU;

/* src/gpudb/KernelArgs.h:306-315 */
typedef struct offsetGPU {
  int* lo_off; int* dim_off1; int* dim_off2; int* dim_off3; int* dim_off4;
} offsetGPU;

/* A NULL struct pointer means "all fields NULL / zero". */

/* ---------------- fused filter + probe + group-by ---------------------- */

/* filter_probe_group_by_GPU2, src/gpudb/GPUProcessing.h:17-155 */
int mordred_filter_probe_group_by_GPU2(const int* gpuCache, const filterArgsGPU* fargs,
    const probeArgsGPU* pargs, const groupbyArgsGPU* gargs, int num_tuples, int* res,
    int start_offset, const short* segment_group, mordred_stream_t stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops and harnesses that forbid
 * hipMalloc): mordred_filter_probe_group_by_GPU2 with the same arguments and
 * results, whose compact probe tables (the per-call pre-pass products) live in
 * a fixed per-GPU scratch array of MORDRED_NOALLOC_SCRATCH_BYTES, held until
 * the process ends, instead of stream-ordered pool memory.  It calls no
 * hipMalloc, hipFree or pool allocation.  A query whose compact tables need
 * more than the array (about 2.1 bytes per slot of each compacted hash table,
 * plus 8.5 KiB per table) is refused with hipErrorOutOfMemory (2) before
 * anything is launched.  Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU (different streams or threads) race. */
#define MORDRED_NOALLOC_SCRATCH_BYTES (16u << 20)
int mordred_filter_probe_group_by_GPU2_noalloc(const int* gpuCache, const filterArgsGPU* fargs,
    const probeArgsGPU* pargs, const groupbyArgsGPU* gargs, int num_tuples, int* res,
    int start_offset, const short* segment_group, mordred_stream_t stream);
/* filter_probe_group_by_GPU3, GPUProcessing.h:157-272 */
int mordred_filter_probe_group_by_GPU3(const int* gpuCache, const offsetGPU* offset,
    const filterArgsGPU* fargs, const probeArgsGPU* pargs, const groupbyArgsGPU* gargs,
    int num_tuples, int* res, mordred_stream_t stream);
/* probe_group_by_GPU2, GPUProcessing.h:275-391 */
int mordred_probe_group_by_GPU2(const int* gpuCache, const probeArgsGPU* pargs,
    const groupbyArgsGPU* gargs, int num_tuples, int* res, int start_offset,
    const short* segment_group, mordred_stream_t stream);
/* probe_group_by_GPU3, GPUProcessing.h:393-493 */
int mordred_probe_group_by_GPU3(const int* gpuCache, const offsetGPU* offset,
    const probeArgsGPU* pargs, const groupbyArgsGPU* gargs, int num_tuples, int* res,
    mordred_stream_t stream);
/* groupByGPU, GPUProcessing.h:1385-1470 */
int mordred_groupByGPU(const int* gpuCache, const offsetGPU* offset, const groupbyArgsGPU* gargs,
    int num_tuples, int* res, mordred_stream_t stream);

/* ---------------- compaction: filter / probe -> offset lists ----------- */

/* filter_probe_GPU2, GPUProcessing.h:496-600 */
int mordred_filter_probe_GPU2(const int* gpuCache, const filterArgsGPU* fargs,
    const probeArgsGPU* pargs, const offsetGPU* out_off, int num_tuples, int* total,
    int start_offset, const short* segment_group, mordred_stream_t stream);
/* filter_probe_GPU3, GPUProcessing.h:602-688 */
int mordred_filter_probe_GPU3(const int* gpuCache, const offsetGPU* in_off,
    const filterArgsGPU* fargs, const probeArgsGPU* pargs, const offsetGPU* out_off,
    int num_tuples, int* total, mordred_stream_t stream);
/* probe_GPU2, GPUProcessing.h:690-814 */
int mordred_probe_GPU2(const int* gpuCache, const probeArgsGPU* pargs, const offsetGPU* out_off,
    int num_tuples, int* total, int start_offset, const short* segment_group,
    mordred_stream_t stream);
/* probe_GPU3, GPUProcessing.h:817-924 */
int mordred_probe_GPU3(const int* gpuCache, const offsetGPU* in_off, const probeArgsGPU* pargs,
    const offsetGPU* out_off, int num_tuples, int* total, mordred_stream_t stream);
/* filter_GPU2, GPUProcessing.h:1174-1286 */
int mordred_filter_GPU2(const int* gpuCache, const filterArgsGPU* fargs, int* out_off,
    int num_tuples, int* total, int start_offset, const short* segment_group,
    mordred_stream_t stream);
/* filter_GPU3, GPUProcessing.h:1288-1383 */
int mordred_filter_GPU3(const int* gpuCache, const int* off_col, const filterArgsGPU* fargs,
    int* out_off, int num_tuples, int* total, mordred_stream_t stream);

/* ---------------- perfect-hash build ----------------------------------- */

/* build_GPU2, GPUProcessing.h:927-997 */
int mordred_build_GPU2(const int* gpuCache, const filterArgsGPU* fargs, const buildArgsGPU* bargs,
    int num_tuples, int* hash_table, int start_offset, const short* segment_group,
    mordred_stream_t stream);
/* build_GPU3, GPUProcessing.h:999-1028 */
int mordred_build_GPU3(const int* gpuCache, const int* dim_off, const filterArgsGPU* fargs,
    const buildArgsGPU* bargs, int num_tuples, int* hash_table, mordred_stream_t stream);

/* build_GPU_minmax, GPUProcessing.h:1030-1113: build_GPU2 plus the min and
 * max of the keys it inserts, merged with atomicMin / atomicMax into the
 * caller-initialised DEVICE ints *min_global / *max_global.  Each workgroup
 * starts from (bargs->val_max, bargs->val_min), as the reference's threads do
 * (:1050), so those two values take part in the merge whenever num_tuples > 0.
 * The reference engine never launches it. */
int mordred_build_GPU_minmax(const int* gpuCache, const filterArgsGPU* fargs, const buildArgsGPU* bargs,
    int num_tuples, int* hash_table, int* min_global, int* max_global, int start_offset,
    const short* segment_group, mordred_stream_t stream);
/* build_GPU_minmax2, GPUProcessing.h:1115-1172: the offset-list twin (rows
 * from the DEVICE list dim_off, no filter).  The keys are those of the listed
 * rows, as build_CPU_minmax2 (CPUProcessing.cu:1910-1983) reads them; the
 * reference's unguarded BlockMinMaxGPU2 overload passes the wrong arguments
 * (BlockLibrary.cuh:1011-1015), which is why the kernel is dead there. */
int mordred_build_GPU_minmax2(const int* gpuCache, const int* dim_off, const filterArgsGPU* fargs,
    const buildArgsGPU* bargs, int num_tuples, int* hash_table, int* min_global, int* max_global,
    mordred_stream_t stream);

/* ---------------- scalar aggregates ------------------------------------ */

/* aggregationGPU, GPUProcessing.h:1472-1523 */
int mordred_aggregationGPU(const int* gpuCache, const int* lo_off, const groupbyArgsGPU* gargs,
    int num_tuples, int* res, mordred_stream_t stream);
/* probe_aggr_GPU2, GPUProcessing.h:1526-1608 */
int mordred_probe_aggr_GPU2(const int* gpuCache, const probeArgsGPU* pargs,
    const groupbyArgsGPU* gargs, int num_tuples, int* res, int start_offset,
    const short* segment_group, mordred_stream_t stream);
/* probe_aggr_GPU3, GPUProcessing.h:1610-1672 */
int mordred_probe_aggr_GPU3(const int* gpuCache, const offsetGPU* offset, const probeArgsGPU* pargs,
    const groupbyArgsGPU* gargs, int num_tuples, int* res, mordred_stream_t stream);
/* filter_probe_aggr_GPU2, GPUProcessing.h:1675-1776 */
int mordred_filter_probe_aggr_GPU2(const int* gpuCache, const filterArgsGPU* fargs,
    const probeArgsGPU* pargs, const groupbyArgsGPU* gargs, int num_tuples, int* res,
    int start_offset, const short* segment_group, mordred_stream_t stream);
/* filter_probe_aggr_GPU3, GPUProcessing.h:1778-1846 */
int mordred_filter_probe_aggr_GPU3(const int* gpuCache, const offsetGPU* offset,
    const filterArgsGPU* fargs, const probeArgsGPU* pargs, const groupbyArgsGPU* gargs,
    int num_tuples, int* res, mordred_stream_t stream);

/* ---------------- on-demand path: plain column pointers ---------------- */

/* build_GPU, GPUProcessing.h:1848-1887.  The reference's device filter
 * function pointer becomes filter_mode (0 EQ, 1 BETWEEN, 2 EQ-or-EQ). */
int mordred_build_GPU(const int* key, const int* val, const int* filter, int compare1,
    int compare2, int filter_mode, int num_tuples, int* hash_table, int num_slots, int val_min,
    int segment_index, mordred_stream_t stream);
/* filter_probe_aggr_GPU, GPUProcessing.h:1890-1964 */
int mordred_filter_probe_aggr_GPU(const int* filter1, const int* filter2, int compare1,
    int compare2, int compare3, int compare4, int filter_mode1, int filter_mode2, const int* key4,
    const int* ht4, int dim_len4, int min_key4, const int* aggr1, const int* aggr2,
    int num_tuples, int* res, mordred_stream_t stream);
/* probe_group_by_GPU, GPUProcessing.h:1966-2055 */
int mordred_probe_group_by_GPU(const int* key1, const int* key2, const int* key3, const int* key4,
    const int* ht1, int dim_len1, const int* ht2, int dim_len2, const int* ht3, int dim_len3,
    const int* ht4, int dim_len4, int min_key1, int min_key2, int min_key3, int min_key4,
    const int* aggr1, const int* aggr2, int min_val1, int min_val2, int min_val3, int min_val4,
    int unique_val1, int unique_val2, int unique_val3, int unique_val4, int total_val,
    int num_tuples, int* res, mordred_stream_t stream);

/* ---------------- registered probe tables ------------------------------ */

/* No counterpart in the reference, whose fused kernels probe each hash table
 * in place.  A fused call derives, per call, a compact stand-in for each hash
 * table it probes often enough (an exact packed form of the slots' results,
 * plus an existence bitmap for large tables).  A caller that probes one table
 * in many calls -- segment groups, streamed chunks, the stages of one query --
 * registers it: both compact forms (value probes and row probes) then get
 * memory now (hipMalloc, about 4.3 bytes per slot plus 17 KiB), are derived
 * once, by the first later call that compacts the table, and are reused by
 * every call after it, on any stream (each waits for that derivation).  A
 * build into the table through any mordred_build_* entry point marks the
 * forms stale, so the next call derives them again; a caller that changes
 * the table in any other way (a memset, a copy) calls
 * mordred_probe_table_invalidate.  Unregister a table only when no call that
 * probes it is still running, as before freeing the table itself.  Answers
 * are the same with or without registration.  Registration is per device
 * (the current one); registering a registered table again with the same
 * dim_len does nothing, with another dim_len re-sizes it. */
int mordred_probe_table_register(const int* hash_table, int dim_len);
int mordred_probe_table_unregister(const int* hash_table);
int mordred_probe_table_invalidate(const int* hash_table);

/* ---------------- library info ----------------------------------------- */

/* Version string of this library. */
const char* mordred_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MORDRED_HIP_H */
