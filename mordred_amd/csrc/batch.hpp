// Shared-scan execution of a batch of star queries (up to 64) on gfx950.
//
// One persistent kernel scans the fact columns once for the whole batch.  Per
// fact row it evaluates every query's fact predicates as bits of a 64-bit
// "alive" word, probes each joined dimension ONCE through a per-dimension
// record table, and accumulates every (row, query) pair that survives.
//
// Record table of dimension d (one 16-byte record per hash slot, HASH =
// (key-min) % len, the slot of the reference's perfect-hash table,
// crystal/join.cuh:3):
//   x, y  64-bit mask: bit q set iff query q joins d and the dimension row in
//         this slot passes q's dimension filter (the row query q's own
//         build_GPU2 would have inserted, GPUProcessing.h:927-997);
//   z     two packed 16-bit group attributes of the row (value - base);
//   w     1 when a row occupies the slot (collision check).
// A fact row survives for query q iff every table q joins has bit q set in the
// probed slot -- exactly the rows q's own fused kernel keeps -- and its group
// value for table d is the attribute q groups on (the hash-table value word
// of q's own build), so every query's result equals its plain run bit for bit.
//
// Everything that changes between batches lives in a device BatchDesc, so
// one captured hipGraph (fill, record builds, scan) serves every batch of the
// same dataset and launch configuration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mordred {

constexpr int kBatchMaxQ = 64;
constexpr int kBatchMaxCols = 12;   // fact columns; slots 0..3 are the dimension keys
constexpr int kBatchMaxPred = 2 * kBatchMaxQ;
constexpr int kBatchMaxExpr = 4;

enum BatchExprKind { kExprDiff = 0, kExprProduct = 1 };  // a1 - a2 (group-by), a1 * a2 (aggregate)
enum BatchQueryFlags { kQueryAggr = 1 };

struct BatchQuery {            // 64 bytes; staged in LDS by the scan kernel
  long long res;               // device address of the query's 6*total_val ints
  int total_val;
  short expr;                  // index into BatchDesc::expr
  short flags;                 // kQueryAggr: no key words
  int lds;                     // u64 offset of its group array in LDS, -1 = global atomics
  int touch;                   // u32 offset of its LDS touched bitmap (when lds >= 0)
  int gtouch;                  // u32 offset of its touched bitmap in the global arena,
                               // valid whatever the placement (a query of a request
                               // loop may change placement between launches)
  signed char attr[4];         // per dimension slot: -1 no group value, else attribute 0 / 1
  int minv[4];
  int uniq[4];
};
static_assert(sizeof(BatchQuery) == 64, "BatchQuery is staged as 4 x 16 bytes");

struct BatchPred {             // fact BETWEEN predicate of one query
  int col, lo, hi, q;
};

struct BatchExpr {
  int kind, a1, a2, pad;       // a2 < 0: absent (0)
};

struct BatchFilt {             // dimension filter of one query (build_GPU2 modes)
  const int* col;              // segment table of the filter column, NULL = pass
  int mode, c1, c2, q;         // mode 0 EQ, 1 BETWEEN, 2 EQ-or-EQ
};

struct BatchDim {
  const int* key;              // segment table of the dimension key column
  const int* attr[2];          // segment tables of the packed attributes (or NULL)
  int attr_base[2];
  uint4* rec;                  // record table, len entries
  unsigned long long uses;     // queries that probe this dimension
  int used, len, minkey, nrows;
  int nfilt, pad;
  BatchFilt filt[kBatchMaxQ];  // sorted by column
};

struct BatchDesc {
  int nq, ncols, npred, nexpr;
  unsigned long long all;      // queries of the batch (dead ones cleared)
  // row range of this launch: positions [0, num_tuples) of the segment group
  // starting at seg_group[seg_begin] (a chunk of a circular scan, or all)
  int seg_begin, num_tuples;
  int rebuild;                 // record tables must be rebuilt (query set changed)
  int pad0;
  unsigned long long reset;    // queries whose result / touched bits are cleared first
  int lds_sum_words;           // u64 words of LDS group arrays
  int lds_touch_words;         // u32 words of LDS touched bitmaps
  long long res_ints;          // result arena ints
  long long touch_words;       // global touched words
  int* res_arena;
  unsigned* touch_arena;
  const int* col[kBatchMaxCols];
  int pred_begin[kBatchMaxCols + 1];
  BatchPred pred[kBatchMaxPred];
  BatchExpr expr[kBatchMaxExpr];
  BatchDim dim[4];
  BatchQuery q[kBatchMaxQ];
};

struct BatchLaunch {           // kernel arguments, fixed for a captured graph
  const BatchDesc* desc;
  const int* cache;
  const short* seg_group;
  int max_tuples;              // largest num_tuples of any launch (sizes the grid)
  unsigned* status;            // [0] = hash-slot collisions in the record builds
};

// Hot part of the descriptor, staged in LDS once per workgroup so the tile
// loop reads it with broadcast LDS loads instead of dependent global loads.
struct BatchHot {
  const int* col[kBatchMaxCols];
  const uint4* rec[4];
  unsigned long long uses[4];
  int len[4], minkey[4];
  int ncols, colmask, usedmask, nexpr, npred, nq;
  int pred_begin[kBatchMaxCols + 1];
  int4 expr[kBatchMaxExpr];
  int4 pred[kBatchMaxPred];    // lo, hi - lo (unsigned), query, column
};

struct BatchConfig {
  int block;                   // 256, 512, 1024
  int vpt;                     // 1, 2 (items per thread = 4*vpt)
  int lds_bytes;               // LDS per workgroup: staged descriptor, query
                               // descriptors and LDS group arrays
};

// Bytes of LDS a query set needs beyond the descriptors.
inline size_t batch_group_lds(int total_val) {
  return (size_t)total_val * 8 + (size_t)((total_val + 31) / 32) * 4;
}

// Host launchers (batch.hip).  Every launch reads the batch from L.desc.
hipError_t batch_fill(const BatchLaunch& L, const uint4* const* rec, const int* rec_len, hipStream_t s);
hipError_t batch_records(const BatchLaunch& L, int d, int nrows, hipStream_t s);
// ncols: fact columns the batch reads (selects the register layout).
hipError_t batch_scan(const BatchLaunch& L, const BatchConfig& c, int ncols, hipStream_t s);
bool batch_config_ok(const BatchConfig& c);
// Allow c.lds_bytes of dynamic LDS for the scan kernel (call before capture).
hipError_t batch_prepare(const BatchConfig& c, int ncols);
int batch_resident(const BatchConfig& c, int ncols);

}  // namespace mordred
