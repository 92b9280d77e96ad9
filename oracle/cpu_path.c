/*
 * ORACLE -- test infrastructure, never the product.  (See oracle.h.)
 *
 * Plain C restatement of src/gpudb/CPUProcessing.cu, one function per
 * reference function.  Structure follows the reference: the row range of a
 * segment group is cut into TASK_SIZE (1,024-row) tasks run in parallel
 * (OpenMP here, TBB parallel_for in the reference), aggregates are kept as a
 * per-thread int64 partial merged with one atomic add, group-by sums use one
 * atomic 64-bit add per row, and compaction reserves output space with one
 * atomic fetch-add per task (CPUProcessing.cu:78, 1706, 509).
 *
 * Integer semantics: the reference multiplies and subtracts in int32 and then
 * widens (CPUProcessing.cu:505-509, 1673).  Here the int32 operations are done
 * in uint32 so that a value outside the specified domain wraps instead of being
 * undefined; inside the domain (|a1*a2| < 2^31) the results are identical.
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define SEG ORACLE_SEGMENT_SIZE
#define TASK ORACLE_TASK_SIZE

/* crystal/join.cuh:3  HASH(X,Y,Z) = ((X-Z) % Y) */
static inline int hash_slot(int key, int len, int min_key) { return (key - min_key) % len; }

static inline int mul32(int a, int b) { return (int)((uint32_t)a * (uint32_t)b); }
static inline int sub32(int a, int b) { return (int)((uint32_t)a - (uint32_t)b); }

static inline long long slot64(const int* ht, int h) {
  /* little-endian [low word = value, high word = row+1] (BlockLibrary.cuh:98) */
  uint64_t lo = (uint32_t)ht[2 * h];
  uint64_t hi = (uint32_t)ht[2 * h + 1];
  return (long long)(lo | (hi << 32));
}

static inline void atomic_add_u64(int* res, int word, long long v) {
  __atomic_fetch_add((unsigned long long*)&res[word], (unsigned long long)v, __ATOMIC_RELAXED);
}

/* row id of segment-group position i (CPUProcessing.cu:31) */
static inline int seg_row(const short* segment_group, int i) {
  return segment_group[i / SEG] * SEG + (i % SEG);
}

static int g_threads = 0;
void oracle_set_threads(int n) { g_threads = n; }
int oracle_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

#define TASK_LOOP(num_tuples)                                                \
  int task_count = ((num_tuples) + TASK - 1) / TASK;                         \
  _Pragma("omp parallel for schedule(dynamic, 16) num_threads(oracle_get_threads())") \
  for (int task = 0; task < task_count; task++)

#define TASK_RANGE(num_tuples)                                               \
  int t_start = task * TASK;                                                 \
  int t_end = t_start + TASK < (num_tuples) ? t_start + TASK : (num_tuples);

/* ---------------------------------------------------------------- filters */

/* Fact BETWEEN filters of the fused operators: CPUProcessing.cu:36-38 (CPU,
 * both columns required) and GPUProcessing.h:72-86 (GPU, NULL column skipped). */
static inline int fact_filters(const oracle_filterArgs* f, int row, int conv) {
  if (conv == ORACLE_CONV_CPU || f->filter_col1 != NULL)
    if (!(f->filter_col1[row] >= f->compare1 && f->filter_col1[row] <= f->compare2)) return 0;
  if (conv == ORACLE_CONV_CPU || f->filter_col2 != NULL)
    if (!(f->filter_col2[row] >= f->compare3 && f->filter_col2[row] <= f->compare4)) return 0;
  return 1;
}

/* Fact BETWEEN filters with NULL checks on both code bases
 * (CPUProcessing.cu:755-761, GPUProcessing.h:72-86). */
static inline int fact_filters_nullable(const oracle_filterArgs* f, int row) {
  if (f->filter_col1 != NULL)
    if (!(f->filter_col1[row] >= f->compare1 && f->filter_col1[row] <= f->compare2)) return 0;
  if (f->filter_col2 != NULL)
    if (!(f->filter_col2[row] >= f->compare3 && f->filter_col2[row] <= f->compare4)) return 0;
  return 1;
}

/* Mode filters of filter/build: CPUProcessing.cu:1189-1200, 1064-1069
 * (mode 1 BETWEEN, mode 2 EQ-or-EQ, anything else passes) and
 * GPUProcessing.h:1231-1254, 972-980 (mode 0 EQ, 1 BETWEEN, 2 EQ-or-EQ; the
 * second filter is AND-ed with the crystal And/Or predicates). */
static inline int mode_filter1(const oracle_filterArgs* f, int row, int conv) {
  if (f->filter_col1 == NULL) return 1;
  int x = f->filter_col1[row];
  switch (f->mode1) {
    case 0: return conv == ORACLE_CONV_GPU ? (x == f->compare1) : 1;
    case 1: return x >= f->compare1 && x <= f->compare2;
    case 2: return x == f->compare1 || x == f->compare2;
    default: return 1;
  }
}

static inline int mode_filter2(const oracle_filterArgs* f, int row, int flag, int conv) {
  if (f->filter_col2 == NULL) return flag;
  int x = f->filter_col2[row];
  switch (f->mode2) {
    case 0: return conv == ORACLE_CONV_GPU ? (flag && x == f->compare3) : flag;
    case 1: return flag && x >= f->compare3 && x <= f->compare4;
    case 2:
      if (conv == ORACLE_CONV_GPU) return (flag && x == f->compare3) || x == f->compare4;
      return flag && (x == f->compare3 || x == f->compare4);
    default: return flag;
  }
}

/* ---------------------------------------------------------------- probes */

/* Probe returning the dimension row (slot high word - 1).  Returns 0 on a miss.
 * The two code bases test different words of the slot:
 *   CPU  CPUProcessing.cu:203-207   the whole 64-bit slot != 0
 *   GPU  BlockLibrary.cuh:24-30     the row word ht[2h+1] != 0 only
 * They agree whenever every build writes the row word (all builds do). */
static inline int probe_row(const int* key_col, const int* ht, int len, int min_key, int row,
                            int* dim_row, int conv) {
  int h = hash_slot(key_col[row], len, min_key);
  if (conv == ORACLE_CONV_GPU) {
    int r = ht[2 * h + 1];
    if (r == 0) return 0;
    *dim_row = r - 1;
    return 1;
  }
  long long s = slot64(ht, h);
  if (s == 0) return 0;
  *dim_row = (int)((unsigned long long)s >> 32) - 1;
  return 1;
}

/* Probe returning the group value (slot low word): CPUProcessing.cu:468-473,
 * BlockLibrary.cuh:98-104.  Returns 0 on a miss. */
static inline int probe_val(const int* key_col, const int* ht, int len, int min_key, int row,
                            int* val) {
  int h = hash_slot(key_col[row], len, min_key);
  long long s = slot64(ht, h);
  if (s == 0) return 0;
  *val = (int)s;
  return 1;
}

/* --------------------------------------------------------------- group-by */

/* Dense group index and accumulation: CPUProcessing.cu:497-509 (CPU),
 * GPUProcessing.h:142-150 (GPU). */
static inline void group_accumulate(const oracle_groupbyArgs* g, int v1, int v2, int v3, int v4,
                                    int row, int* res, int conv) {
  int h = (sub32(v1, g->min_val1) * g->unique_val1 + sub32(v2, g->min_val2) * g->unique_val2 +
           sub32(v3, g->min_val3) * g->unique_val3 + sub32(v4, g->min_val4) * g->unique_val4) %
          g->total_val;
  if (conv == ORACLE_CONV_GPU) {
    res[h * 6] = v1; res[h * 6 + 1] = v2; res[h * 6 + 2] = v3; res[h * 6 + 3] = v4;
  } else {
    if (v1 != 0) res[h * 6] = v1;
    if (v2 != 0) res[h * 6 + 1] = v2;
    if (v3 != 0) res[h * 6 + 2] = v3;
    if (v4 != 0) res[h * 6 + 3] = v4;
  }
  int a1 = g->aggr_col1 ? g->aggr_col1[row] : 0;
  int a2 = g->aggr_col2 ? g->aggr_col2[row] : 0;
  atomic_add_u64(res, h * 6 + 4, (long long)sub32(a1, a2));
}

/* Probe all four tables for group values, or (offset mode) take the group
 * value from an earlier probe's dimension row.  Returns 0 if a probe misses.
 * CPUProcessing.cu:597-636 and GPUProcessing.h:200-238. */
static int probe_groups(const oracle_probeArgs* p, const oracle_groupbyArgs* g,
                        const oracle_offset* in, int pos, int row, int* v) {
  const int* keys[4] = {p->key_col1, p->key_col2, p->key_col3, p->key_col4};
  const int* hts[4] = {p->ht1, p->ht2, p->ht3, p->ht4};
  const int lens[4] = {p->dim_len1, p->dim_len2, p->dim_len3, p->dim_len4};
  const int mins[4] = {p->min_key1, p->min_key2, p->min_key3, p->min_key4};
  const int* gcols[4] = {g->group_col1, g->group_col2, g->group_col3, g->group_col4};
  int* doffs[4] = {NULL, NULL, NULL, NULL};
  if (in) { doffs[0] = in->h_dim_off1; doffs[1] = in->h_dim_off2; doffs[2] = in->h_dim_off3; doffs[3] = in->h_dim_off4; }
  for (int k = 0; k < 4; k++) {
    v[k] = 0;
    if (keys[k] != NULL && hts[k] != NULL) {
      if (!probe_val(keys[k], hts[k], lens[k], mins[k], row, &v[k])) return 0;
    } else if (in != NULL && gcols[k] != NULL) {
      v[k] = gcols[k][doffs[k][pos]];
    }
  }
  return 1;
}

/* c5: probe_group_by_CPU  CPUProcessing.cu:439-568 */
void oracle_probe_group_by_CPU(oracle_probeArgs pargs, oracle_groupbyArgs gargs, int num_tuples,
                               int* res, int start_offset, const short* segment_group, int conv) {
  (void)start_offset;
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int row = seg_row(segment_group, i), v[4];
      if (!probe_groups(&pargs, &gargs, NULL, i, row, v)) continue;
      group_accumulate(&gargs, v[0], v[1], v[2], v[3], row, res, conv);
    }
  }
}

/* c6: probe_group_by_CPU2  CPUProcessing.cu:570-721 */
void oracle_probe_group_by_CPU2(oracle_offset offset, oracle_probeArgs pargs,
                                oracle_groupbyArgs gargs, int num_tuples, int* res,
                                int start_offset, int conv) {
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int pos = start_offset + i, row = offset.h_lo_off[pos], v[4];
      if (!probe_groups(&pargs, &gargs, &offset, pos, row, v)) continue;
      group_accumulate(&gargs, v[0], v[1], v[2], v[3], row, res, conv);
    }
  }
}

/* c7: filter_probe_group_by_CPU  CPUProcessing.cu:723-866 */
void oracle_filter_probe_group_by_CPU(oracle_filterArgs fargs, oracle_probeArgs pargs,
                                      oracle_groupbyArgs gargs, int num_tuples, int* res,
                                      int start_offset, const short* segment_group, int conv) {
  (void)start_offset;
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int row = seg_row(segment_group, i), v[4];
      if (!fact_filters_nullable(&fargs, row)) continue;
      if (!probe_groups(&pargs, &gargs, NULL, i, row, v)) continue;
      group_accumulate(&gargs, v[0], v[1], v[2], v[3], row, res, conv);
    }
  }
}

/* c8: filter_probe_group_by_CPU2  CPUProcessing.cu:868-1032 */
void oracle_filter_probe_group_by_CPU2(oracle_offset offset, oracle_filterArgs fargs,
                                       oracle_probeArgs pargs, oracle_groupbyArgs gargs,
                                       int num_tuples, int* res, int start_offset, int conv) {
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int pos = start_offset + i, row = offset.h_lo_off[pos], v[4];
      if (!fact_filters_nullable(&fargs, row)) continue;
      if (!probe_groups(&pargs, &gargs, &offset, pos, row, v)) continue;
      group_accumulate(&gargs, v[0], v[1], v[2], v[3], row, res, conv);
    }
  }
}

/* c13: groupByCPU  CPUProcessing.cu:1335-1438 (GPU twin groupByGPU,
 * GPUProcessing.h:1385-1470): every offset-list row counts. */
void oracle_groupByCPU(oracle_offset offset, oracle_groupbyArgs gargs, int num_tuples, int* res,
                       int conv) {
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int v1 = gargs.group_col1 ? gargs.group_col1[offset.h_dim_off1[i]] : 0;
      int v2 = gargs.group_col2 ? gargs.group_col2[offset.h_dim_off2[i]] : 0;
      int v3 = gargs.group_col3 ? gargs.group_col3[offset.h_dim_off3[i]] : 0;
      int v4 = gargs.group_col4 ? gargs.group_col4[offset.h_dim_off4[i]] : 0;
      group_accumulate(&gargs, v1, v2, v3, v4, offset.h_lo_off[i], res, conv);
    }
  }
}

/* ------------------------------------------------------------- compaction */

/* Reserve `count` output rows with one atomic fetch-add (CPUProcessing.cu:78). */
static inline int reserve(int* total, int count) {
  return __atomic_fetch_add(total, count, __ATOMIC_RELAXED);
}

static void emit_offsets(const oracle_offset* out, int base, int count, int* const tmp[5]) {
  int* dst[5] = {out->h_lo_off, out->h_dim_off1, out->h_dim_off2, out->h_dim_off3, out->h_dim_off4};
  for (int c = 0; c < 5; c++)
    if (dst[c] != NULL) memcpy(dst[c] + base, tmp[c], sizeof(int) * (size_t)count);
}

/* Probe the four tables returning dimension rows; unprobed tables keep the
 * incoming offset (offset mode) or the "not probed" value: 0 on the CPU
 * (slot default 1, minus 1; CPUProcessing.cu:197, 229-233) and 1 on the GPU
 * (BlockSetValue(..., 1), GPUProcessing.h:757). */
static int probe_rows(const oracle_probeArgs* p, const oracle_offset* in, int pos, int row,
                      int conv, int* d) {
  const int* keys[4] = {p->key_col1, p->key_col2, p->key_col3, p->key_col4};
  const int* hts[4] = {p->ht1, p->ht2, p->ht3, p->ht4};
  const int lens[4] = {p->dim_len1, p->dim_len2, p->dim_len3, p->dim_len4};
  const int mins[4] = {p->min_key1, p->min_key2, p->min_key3, p->min_key4};
  int* doffs[4] = {NULL, NULL, NULL, NULL};
  if (in) { doffs[0] = in->h_dim_off1; doffs[1] = in->h_dim_off2; doffs[2] = in->h_dim_off3; doffs[3] = in->h_dim_off4; }
  for (int k = 0; k < 4; k++) {
    d[k] = conv == ORACLE_CONV_GPU ? 1 : 0;
    if (keys[k] != NULL && hts[k] != NULL) {
      if (!probe_row(keys[k], hts[k], lens[k], mins[k], row, &d[k], conv)) return 0;
    } else if (doffs[k] != NULL) {
      d[k] = doffs[k][pos];
    }
  }
  return 1;
}

#define COMPACT_BEGIN                                                      \
  int buf0[TASK], buf1[TASK], buf2[TASK], buf3[TASK], buf4[TASK];          \
  int* const tmp[5] = {buf0, buf1, buf2, buf3, buf4};                     \
  int count = 0;

/* c3: probe_CPU  CPUProcessing.cu:170-299 (GPU twin probe_GPU2, GPUProcessing.h:690-814) */
void oracle_probe_CPU(oracle_probeArgs pargs, oracle_offset out_off, int num_tuples, int* total,
                      int start_offset, const short* segment_group, int conv) {
  (void)start_offset;
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    COMPACT_BEGIN
    for (int i = t_start; i < t_end; i++) {
      int row = seg_row(segment_group, i), d[4];
      if (!probe_rows(&pargs, NULL, i, row, conv, d)) continue;
      buf0[count] = row; buf1[count] = d[0]; buf2[count] = d[1]; buf3[count] = d[2]; buf4[count] = d[3];
      count++;
    }
    emit_offsets(&out_off, reserve(total, count), count, tmp);
  }
}

/* c4: probe_CPU2  CPUProcessing.cu:301-437 (GPU twin probe_GPU3, GPUProcessing.h:817-924) */
void oracle_probe_CPU2(oracle_offset in_off, oracle_probeArgs pargs, oracle_offset out_off,
                       int num_tuples, int* total, int start_offset, int conv) {
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    COMPACT_BEGIN
    for (int i = t_start; i < t_end; i++) {
      int pos = start_offset + i, row = in_off.h_lo_off[pos], d[4];
      if (!probe_rows(&pargs, &in_off, pos, row, conv, d)) continue;
      buf0[count] = row; buf1[count] = d[0]; buf2[count] = d[1]; buf3[count] = d[2]; buf4[count] = d[3];
      count++;
    }
    emit_offsets(&out_off, reserve(total, count), count, tmp);
  }
}

/* c1: filter_probe_CPU  CPUProcessing.cu:3-86 (GPU twin filter_probe_GPU2,
 * GPUProcessing.h:496-600): BETWEEN x2, probe table 4 only. */
void oracle_filter_probe_CPU(oracle_filterArgs fargs, oracle_probeArgs pargs,
                             oracle_offset out_off, int num_tuples, int* total, int start_offset,
                             const short* segment_group, int conv) {
  (void)start_offset;
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    COMPACT_BEGIN
    for (int i = t_start; i < t_end; i++) {
      int row = seg_row(segment_group, i);
      int d4 = conv == ORACLE_CONV_GPU ? 1 : 0;
      if (!fact_filters(&fargs, row, conv)) continue;
      if (conv == ORACLE_CONV_CPU || (pargs.key_col4 != NULL && pargs.ht4 != NULL))
        if (!probe_row(pargs.key_col4, pargs.ht4, pargs.dim_len4, pargs.min_key4, row, &d4, conv)) continue;
      buf0[count] = row; buf4[count] = d4;
      count++;
    }
    oracle_offset o = out_off;
    o.h_dim_off1 = o.h_dim_off2 = o.h_dim_off3 = NULL;
    emit_offsets(&o, reserve(total, count), count, tmp);
  }
}

/* c2: filter_probe_CPU2  CPUProcessing.cu:88-168 (GPU twin filter_probe_GPU3,
 * GPUProcessing.h:602-688; the GPU passes an incoming dim_off4 through when
 * table 4 is not probed). */
void oracle_filter_probe_CPU2(oracle_offset in_off, oracle_filterArgs fargs,
                              oracle_probeArgs pargs, oracle_offset out_off, int num_tuples,
                              int* total, int start_offset, int conv) {
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    COMPACT_BEGIN
    for (int i = t_start; i < t_end; i++) {
      int pos = start_offset + i, row = in_off.h_lo_off[pos];
      int d4 = conv == ORACLE_CONV_GPU ? 1 : 0;
      if (!fact_filters(&fargs, row, conv)) continue;
      if (conv == ORACLE_CONV_CPU || (pargs.key_col4 != NULL && pargs.ht4 != NULL)) {
        if (!probe_row(pargs.key_col4, pargs.ht4, pargs.dim_len4, pargs.min_key4, row, &d4, conv)) continue;
      } else if (in_off.h_dim_off4 != NULL) {
        d4 = in_off.h_dim_off4[pos];
      }
      buf0[count] = row; buf4[count] = d4;
      count++;
    }
    oracle_offset o = out_off;
    o.h_dim_off1 = o.h_dim_off2 = o.h_dim_off3 = NULL;
    emit_offsets(&o, reserve(total, count), count, tmp);
  }
}

/* c11: filter_CPU  CPUProcessing.cu:1158-1246 (GPU twin filter_GPU2, GPUProcessing.h:1174-1286) */
void oracle_filter_CPU(oracle_filterArgs fargs, int* out_off, int num_tuples, int* total,
                       int start_offset, const short* segment_group, int conv) {
  (void)start_offset;
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    COMPACT_BEGIN
    for (int i = t_start; i < t_end; i++) {
      int row = seg_row(segment_group, i);
      int flag = mode_filter2(&fargs, row, mode_filter1(&fargs, row, conv), conv);
      if (flag) buf0[count++] = row;
    }
    oracle_offset o = {out_off, NULL, NULL, NULL, NULL};
    emit_offsets(&o, reserve(total, count), count, tmp);
  }
}

/* c12: filter_CPU2  CPUProcessing.cu:1248-1333 (GPU twin filter_GPU3, GPUProcessing.h:1288-1383) */
void oracle_filter_CPU2(const int* off_col, oracle_filterArgs fargs, int* out_off, int num_tuples,
                        int* total, int start_offset, int conv) {
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    COMPACT_BEGIN
    for (int i = t_start; i < t_end; i++) {
      int row = off_col[start_offset + i];
      int flag = mode_filter2(&fargs, row, mode_filter1(&fargs, row, conv), conv);
      if (flag) buf0[count++] = row;
    }
    oracle_offset o = {out_off, NULL, NULL, NULL, NULL};
    emit_offsets(&o, reserve(total, count), count, tmp);
  }
}

/* ------------------------------------------------------------------ build */

/* Slot write of a qualifying dimension row: CPUProcessing.cu:1073-1075 (plain
 * stores) and BlockLibrary.cuh:613-614 (CAS on the value word, then store). */
static inline void build_slot(int* ht, int h, int row, const int* val_col, int conv) {
  if (val_col != NULL) {
    if (conv == ORACLE_CONV_GPU)
      __atomic_compare_exchange_n(&ht[h << 1], &(int){0}, val_col[row], 0, __ATOMIC_RELAXED,
                                  __ATOMIC_RELAXED);
    else
      ht[h << 1] = val_col[row];
  }
  ht[(h << 1) + 1] = row + 1;
}

/* c9: build_CPU  CPUProcessing.cu:1034-1104 (GPU twin build_GPU2, GPUProcessing.h:927-997) */
void oracle_build_CPU(oracle_filterArgs fargs, oracle_buildArgs bargs, int num_tuples,
                      int* hash_table, int start_offset, const short* segment_group, int conv) {
  (void)start_offset;
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int row = seg_row(segment_group, i);
      if (!mode_filter1(&fargs, row, conv)) continue;
      build_slot(hash_table, hash_slot(bargs.key_col[row], bargs.num_slots, bargs.val_min), row,
                 bargs.val_col, conv);
    }
  }
}

/* c10: build_CPU2  CPUProcessing.cu:1106-1155 (GPU twin build_GPU3, GPUProcessing.h:999-1028):
 * no filter, rows come from a dimension offset list. */
void oracle_build_CPU2(const int* dim_off, oracle_filterArgs fargs, oracle_buildArgs bargs,
                       int num_tuples, int* hash_table, int start_offset, int conv) {
  (void)fargs;
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int row = dim_off[start_offset + i];
      build_slot(hash_table, hash_slot(bargs.key_col[row], bargs.num_slots, bargs.val_min), row,
                 bargs.val_col, conv);
    }
  }
}

static void minmax_merge(int* min_global, int* max_global, int mn, int mx) {
  int prev = __atomic_load_n(max_global, __ATOMIC_RELAXED);
  while (mx > prev && !__atomic_compare_exchange_n(max_global, &prev, mx, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
  prev = __atomic_load_n(min_global, __ATOMIC_RELAXED);
  while (mn < prev && !__atomic_compare_exchange_n(min_global, &prev, mn, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
}

/* c20: build_CPU_minmax  CPUProcessing.cu:1821-1908: build + min/max of the
 * built keys; per-thread extremes start at (val_max, val_min). */
void oracle_build_CPU_minmax(oracle_filterArgs fargs, oracle_buildArgs bargs, int num_tuples,
                             int* hash_table, int* min_global, int* max_global, int start_offset,
                             const short* segment_group, int conv) {
  (void)start_offset;
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    int mn = bargs.val_max, mx = bargs.val_min;
    for (int i = t_start; i < t_end; i++) {
      int row = seg_row(segment_group, i);
      if (!mode_filter1(&fargs, row, conv)) continue;
      int key = bargs.key_col[row];
      if (key < mn) mn = key;
      if (key > mx) mx = key;
      build_slot(hash_table, hash_slot(key, bargs.num_slots, bargs.val_min), row, bargs.val_col, conv);
    }
    minmax_merge(min_global, max_global, mn, mx);
  }
}

/* c21: build_CPU_minmax2  CPUProcessing.cu:1910-1983 */
void oracle_build_CPU_minmax2(const int* dim_off, oracle_filterArgs fargs, oracle_buildArgs bargs,
                              int num_tuples, int* hash_table, int* min_global, int* max_global,
                              int start_offset, int conv) {
  (void)fargs;
  TASK_LOOP(num_tuples) {
    TASK_RANGE(num_tuples)
    int mn = bargs.val_max, mx = bargs.val_min;
    for (int i = t_start; i < t_end; i++) {
      int row = dim_off[start_offset + i];
      int key = bargs.key_col[row];
      if (key < mn) mn = key;
      if (key > mx) mx = key;
      build_slot(hash_table, hash_slot(key, bargs.num_slots, bargs.val_min), row, bargs.val_col, conv);
    }
    minmax_merge(min_global, max_global, mn, mx);
  }
}

/* ------------------------------------------------------------ aggregation */

static inline long long aggr_product(const oracle_groupbyArgs* g, int row) {
  int a1 = g->aggr_col1 ? g->aggr_col1[row] : 0;
  int a2 = g->aggr_col2 ? g->aggr_col2[row] : 0;
  return (long long)mul32(a1, a2);
}

/* c14: aggregationCPU  CPUProcessing.cu:1440-1488 */
void oracle_aggregationCPU(const int* lo_off, oracle_groupbyArgs gargs, int num_tuples, int* res) {
  long long sum = 0;
  int task_count = (num_tuples + TASK - 1) / TASK;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : sum) num_threads(oracle_get_threads())
  for (int task = 0; task < task_count; task++) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) sum += aggr_product(&gargs, lo_off[i]);
  }
  atomic_add_u64(res, 4, sum);
}

/* c15: probe_aggr_CPU  CPUProcessing.cu:1491-1558 */
void oracle_probe_aggr_CPU(oracle_probeArgs pargs, oracle_groupbyArgs gargs, int num_tuples,
                           int* res, int start_offset, const short* segment_group) {
  (void)start_offset;
  long long sum = 0;
  int task_count = (num_tuples + TASK - 1) / TASK;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : sum) num_threads(oracle_get_threads())
  for (int task = 0; task < task_count; task++) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int row = seg_row(segment_group, i), v;
      if (pargs.key_col4 != NULL && pargs.ht4 != NULL &&
          !probe_val(pargs.key_col4, pargs.ht4, pargs.dim_len4, pargs.min_key4, row, &v))
        continue;
      sum += aggr_product(&gargs, row);
    }
  }
  atomic_add_u64(res, 4, sum);
}

/* c16: probe_aggr_CPU2  CPUProcessing.cu:1560-1626 */
void oracle_probe_aggr_CPU2(oracle_offset offset, oracle_probeArgs pargs, oracle_groupbyArgs gargs,
                            int num_tuples, int* res, int start_offset) {
  long long sum = 0;
  int task_count = (num_tuples + TASK - 1) / TASK;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : sum) num_threads(oracle_get_threads())
  for (int task = 0; task < task_count; task++) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int row = offset.h_lo_off[start_offset + i], v;
      if (pargs.key_col4 != NULL && pargs.ht4 != NULL &&
          !probe_val(pargs.key_col4, pargs.ht4, pargs.dim_len4, pargs.min_key4, row, &v))
        continue;
      sum += aggr_product(&gargs, row);
    }
  }
  atomic_add_u64(res, 4, sum);
}

/* c17: filter_probe_aggr_CPU  CPUProcessing.cu:1628-1710 (the SSB Q1.1 CPU path;
 * GPU twin filter_probe_aggr_GPU2, GPUProcessing.h:1675-1776) */
void oracle_filter_probe_aggr_CPU(oracle_filterArgs fargs, oracle_probeArgs pargs,
                                  oracle_groupbyArgs gargs, int num_tuples, int* res,
                                  int start_offset, const short* segment_group, int conv) {
  (void)start_offset;
  long long sum = 0;
  int task_count = (num_tuples + TASK - 1) / TASK;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : sum) num_threads(oracle_get_threads())
  for (int task = 0; task < task_count; task++) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int row = seg_row(segment_group, i), v;
      if (!fact_filters(&fargs, row, conv)) continue;
      if ((conv == ORACLE_CONV_CPU || (pargs.key_col4 != NULL && pargs.ht4 != NULL)) &&
          !probe_val(pargs.key_col4, pargs.ht4, pargs.dim_len4, pargs.min_key4, row, &v))
        continue;
      sum += aggr_product(&gargs, row);
    }
  }
  atomic_add_u64(res, 4, sum);
}

/* c18: filter_probe_aggr_CPU2  CPUProcessing.cu:1712-1784 (GPU twin
 * filter_probe_aggr_GPU3, GPUProcessing.h:1778-1846, which re-checks only the
 * second filter) */
void oracle_filter_probe_aggr_CPU2(oracle_offset offset, oracle_filterArgs fargs,
                                   oracle_probeArgs pargs, oracle_groupbyArgs gargs,
                                   int num_tuples, int* res, int start_offset, int conv) {
  long long sum = 0;
  int task_count = (num_tuples + TASK - 1) / TASK;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : sum) num_threads(oracle_get_threads())
  for (int task = 0; task < task_count; task++) {
    TASK_RANGE(num_tuples)
    for (int i = t_start; i < t_end; i++) {
      int row = offset.h_lo_off[start_offset + i], v;
      if (conv == ORACLE_CONV_GPU) {
        if (fargs.filter_col2 != NULL &&
            !(fargs.filter_col2[row] >= fargs.compare3 && fargs.filter_col2[row] <= fargs.compare4))
          continue;
      } else if (!fact_filters(&fargs, row, conv)) {
        continue;
      }
      if ((conv == ORACLE_CONV_CPU || (pargs.key_col4 != NULL && pargs.ht4 != NULL)) &&
          !probe_val(pargs.key_col4, pargs.ht4, pargs.dim_len4, pargs.min_key4, row, &v))
        continue;
      sum += aggr_product(&gargs, row);
    }
  }
  atomic_add_u64(res, 4, sum);
}

/* c19: merge  CPUProcessing.cu:1786-1819: a CPU key word is kept unless it is
 * 0, then the GPU word is taken; the int64 sums are added. */
void oracle_merge(int* resCPU, const int* resGPU, int num_tuples) {
#pragma omp parallel for schedule(static) num_threads(oracle_get_threads())
  for (int i = 0; i < num_tuples; i++) {
    for (int k = 0; k < 4; k++)
      if (resCPU[i * 6 + k] == 0) resCPU[i * 6 + k] = resGPU[i * 6 + k];
    ((unsigned long long*)resCPU)[i * 3 + 2] += ((const unsigned long long*)resGPU)[i * 3 + 2];
  }
}
