/*
 * ORACLE -- test infrastructure, never the product.
 *
 * A plain C restatement of the Mordred engine's CPU query path
 * (src/gpudb/CPUProcessing.cu) used only as the checker for the HIP kernels
 * in mordred_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every function mirrors one reference function (file:line given at each
 * definition in cpu_path.c).  Columns are full host columns indexed by row
 * id, exactly like the reference CPU path: a row of segment-group position i
 * lives at  segment_group[i / SEGMENT_SIZE] * SEGMENT_SIZE + i % SEGMENT_SIZE.
 *
 * The reference has two code bases for the same operators (the CPU functions
 * and the GPU kernels of src/gpudb/GPUProcessing.h) and they differ in a few
 * places.  Every function therefore takes `conv`:
 *   ORACLE_CONV_CPU  follows src/gpudb/CPUProcessing.cu,
 *   ORACLE_CONV_GPU  follows src/gpudb/GPUProcessing.h + BlockLibrary.cuh.
 * The differences (all listed where they are implemented):
 *   - group-by key words: GPU writes all four keys of a selected row,
 *     CPU writes only non-zero keys            (GPUProcessing.h:144-147 vs
 *                                              CPUProcessing.cu:498-501)
 *   - compaction dim_off of a dimension that was not probed: GPU 1, CPU 0
 *                                              (GPUProcessing.h:757 vs
 *                                              CPUProcessing.cu:197,229-233)
 *   - filter / build mode 0: GPU "==", CPU passes every row
 *                                              (GPUProcessing.h:972,1231 vs
 *                                              CPUProcessing.cu:1064,1189)
 *   - second filter in mode 2: GPU (f && x==c3) || x==c4 (crystal OR-pred),
 *     CPU f && (x==c3 || x==c4)                (GPUProcessing.h:1252 vs
 *                                              CPUProcessing.cu:1197)
 *   - offset-mode filter+probe+aggregate: GPU re-checks filter 2 only,
 *     CPU re-checks both                       (GPUProcessing.h:1805 vs
 *                                              CPUProcessing.cu:1730-1736)
 *   - fact filters of the fused kernels: GPU skips a NULL filter column,
 *     the Q1-only CPU functions require both   (GPUProcessing.h:1718 vs
 *                                              CPUProcessing.cu:36-38)
 *   - build value slot: GPU compare-and-swap (first writer wins), CPU plain
 *     store (last writer wins); identical for unique dimension keys.
 *   - row probes (dimension row out): GPU tests the row word ht[2h+1] only,
 *     CPU the whole 64-bit slot            (BlockLibrary.cuh:24-30 vs
 *                                              CPUProcessing.cu:203-207)
 */
#ifndef MORDRED_ORACLE_H
#define MORDRED_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_SEGMENT_SIZE 1048576   /* src/gpudb/common.h:73 */
#define ORACLE_TASK_SIZE 1024         /* src/gpudb/CPUProcessing.h:9 */

enum { ORACLE_CONV_CPU = 0, ORACLE_CONV_GPU = 1 };

/* src/gpudb/KernelArgs.h:288-304 */
typedef struct {
  const int* filter_col1;
  const int* filter_col2;
  int compare1, compare2, compare3, compare4;
  int mode1, mode2;
  void* h_filter_func1;   /* unused: the modes carry the predicate */
  void* h_filter_func2;
} oracle_filterArgs;

/* src/gpudb/KernelArgs.h:171-194 */
typedef struct {
  const int* key_col1; const int* key_col2; const int* key_col3; const int* key_col4;
  const int* ht1; const int* ht2; const int* ht3; const int* ht4;
  int dim_len1, dim_len2, dim_len3, dim_len4;
  int min_key1, min_key2, min_key3, min_key4;
} oracle_probeArgs;

/* src/gpudb/KernelArgs.h:207-216 */
typedef struct {
  const int* key_col;
  const int* val_col;
  int num_slots, val_min, val_max;
} oracle_buildArgs;

/* src/gpudb/KernelArgs.h:244-268 */
typedef struct {
  const int* aggr_col1; const int* aggr_col2;
  const int* group_col1; const int* group_col2; const int* group_col3; const int* group_col4;
  int min_val1, min_val2, min_val3, min_val4;
  int unique_val1, unique_val2, unique_val3, unique_val4;
  int total_val, mode;
  void* h_group_func;     /* unused: the reference hard-codes the operator */
} oracle_groupbyArgs;

/* src/gpudb/KernelArgs.h:317-326 */
typedef struct {
  int* h_lo_off; int* h_dim_off1; int* h_dim_off2; int* h_dim_off3; int* h_dim_off4;
} oracle_offset;

void oracle_set_threads(int n);
int oracle_get_threads(void);

void oracle_filter_probe_CPU(oracle_filterArgs fargs, oracle_probeArgs pargs, oracle_offset out_off,
    int num_tuples, int* total, int start_offset, const short* segment_group, int conv);
void oracle_filter_probe_CPU2(oracle_offset in_off, oracle_filterArgs fargs, oracle_probeArgs pargs,
    oracle_offset out_off, int num_tuples, int* total, int start_offset, int conv);
void oracle_probe_CPU(oracle_probeArgs pargs, oracle_offset out_off, int num_tuples, int* total,
    int start_offset, const short* segment_group, int conv);
void oracle_probe_CPU2(oracle_offset in_off, oracle_probeArgs pargs, oracle_offset out_off,
    int num_tuples, int* total, int start_offset, int conv);
void oracle_probe_group_by_CPU(oracle_probeArgs pargs, oracle_groupbyArgs gargs, int num_tuples,
    int* res, int start_offset, const short* segment_group, int conv);
void oracle_probe_group_by_CPU2(oracle_offset offset, oracle_probeArgs pargs, oracle_groupbyArgs gargs,
    int num_tuples, int* res, int start_offset, int conv);
void oracle_filter_probe_group_by_CPU(oracle_filterArgs fargs, oracle_probeArgs pargs,
    oracle_groupbyArgs gargs, int num_tuples, int* res, int start_offset,
    const short* segment_group, int conv);
void oracle_filter_probe_group_by_CPU2(oracle_offset offset, oracle_filterArgs fargs,
    oracle_probeArgs pargs, oracle_groupbyArgs gargs, int num_tuples, int* res, int start_offset,
    int conv);
void oracle_build_CPU(oracle_filterArgs fargs, oracle_buildArgs bargs, int num_tuples,
    int* hash_table, int start_offset, const short* segment_group, int conv);
void oracle_build_CPU2(const int* dim_off, oracle_filterArgs fargs, oracle_buildArgs bargs,
    int num_tuples, int* hash_table, int start_offset, int conv);
void oracle_filter_CPU(oracle_filterArgs fargs, int* out_off, int num_tuples, int* total,
    int start_offset, const short* segment_group, int conv);
void oracle_filter_CPU2(const int* off_col, oracle_filterArgs fargs, int* out_off, int num_tuples,
    int* total, int start_offset, int conv);
void oracle_groupByCPU(oracle_offset offset, oracle_groupbyArgs gargs, int num_tuples, int* res,
    int conv);
void oracle_aggregationCPU(const int* lo_off, oracle_groupbyArgs gargs, int num_tuples, int* res);
void oracle_probe_aggr_CPU(oracle_probeArgs pargs, oracle_groupbyArgs gargs, int num_tuples,
    int* res, int start_offset, const short* segment_group);
void oracle_probe_aggr_CPU2(oracle_offset offset, oracle_probeArgs pargs, oracle_groupbyArgs gargs,
    int num_tuples, int* res, int start_offset);
void oracle_filter_probe_aggr_CPU(oracle_filterArgs fargs, oracle_probeArgs pargs,
    oracle_groupbyArgs gargs, int num_tuples, int* res, int start_offset,
    const short* segment_group, int conv);
void oracle_filter_probe_aggr_CPU2(oracle_offset offset, oracle_filterArgs fargs,
    oracle_probeArgs pargs, oracle_groupbyArgs gargs, int num_tuples, int* res, int start_offset,
    int conv);
void oracle_merge(int* resCPU, const int* resGPU, int num_tuples);
void oracle_build_CPU_minmax(oracle_filterArgs fargs, oracle_buildArgs bargs, int num_tuples,
    int* hash_table, int* min_global, int* max_global, int start_offset,
    const short* segment_group, int conv);
void oracle_build_CPU_minmax2(const int* dim_off, oracle_filterArgs fargs, oracle_buildArgs bargs,
    int num_tuples, int* hash_table, int* min_global, int* max_global, int start_offset, int conv);

#ifdef __cplusplus
}
#endif

#endif
