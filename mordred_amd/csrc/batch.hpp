// Shared-scan execution of a batch of star queries (up to 64) on gfx950.
//
// One persistent kernel scans the fact columns once for the whole batch.  Per
// fact row it evaluates every query's fact predicates as bits of a 64-bit
// "alive" word, probes each joined dimension ONCE through a packed attribute
// table, and accumulates every (row, query) pair that survives.
//
// Packed table of dimension d (one 1-, 2- or 4-byte entry per hash slot, HASH
// = (key-min) % len, the slot of the reference's perfect-hash table,
// crystal/join.cuh:3): the top bit says a dimension row occupies the slot,
// the other bits hold the row's attributes that the batch filters or groups
// on (value - base).  The mask of queries whose dimension filter the row
// passes is rebuilt per fact row:
//   * a filter on an attribute with a small value span: a 64-bit mask per
//     attribute value (bit q = query q passes, or does not filter on it),
//     staged in LDS and AND-ed over the attributes;
//   * a filter on the dimension key or on a wide attribute: a range compare.
// A fact row survives for query q iff every table q joins holds a row in the
// probed slot that passes q's filter -- exactly the rows q's own build_GPU2
// inserts (GPUProcessing.h:927-997) -- and its group value for table d is the
// attribute q groups on (the value word of q's own hash table), so every
// query's result equals its plain run bit for bit.  The host checks that keys
// cannot collide (dimension and fact key ranges within dim_len), which makes
// the dimension key of a probed row the fact row's key.
//
// Everything that changes between batches lives in a device BatchDesc, so
// one captured hipGraph (fill, packed-table builds, scan) serves every batch
// of the same dataset and launch configuration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mordred {

constexpr int kBatchMaxQ = 64;
constexpr int kBatchMaxCols = 12;   // fact columns; slots 0..3 are the dimension keys
constexpr int kBatchMaxPred = 2 * kBatchMaxQ;
constexpr int kBatchMaxExpr = 4;
constexpr int kBatchMaxAttr = 3;    // packed attributes per dimension
constexpr int kBatchMaxCmp = 64;    // range-compare filters per dimension (one per query)
constexpr int kBatchMaxLut = 4096;  // 64-bit mask entries staged in LDS (32 KiB)
constexpr int kBatchLutSpan = 2048; // widest attribute span given a mask table
constexpr int kBatchIvlMin = 4;     // range filters on one source that get an interval table

enum BatchExprKind { kExprDiff = 0, kExprProduct = 1 };  // a1 - a2 (group-by), a1 * a2 (aggregate)
enum BatchQueryFlags {
  kQueryAggr = 1,        // scalar aggregate: no key words
  kQueryKeysInLoop = 2,  // keys do not decode from the group index: stored by the pair
                         // whose add returns a zero sum (returning atomics)
};
// Where a query's sums accumulate during the scan (per workgroup, in LDS, then
// added to the result once at the end; or per pair in the result itself).
enum BatchSink {
  kSinkLds64 = 0,   // int64 per group in LDS, non-returning adds
  kSinkLds32 = 1,   // low 32 bits per group in LDS; a carry or borrow out of
                    // them adds +-2^32 to the result directly (exact mod 2^64)
  kSinkGlobal = 2,  // one 64-bit atomic per pair on the result
};

// The fields a (row, query) pair reads sit in the first 48 bytes (three
// 16-byte LDS reads).
struct BatchQuery {            // 64 bytes; staged in LDS by the scan kernel
  long long res;               // device address of the query's 6*total_val ints
  int total_val;
  // even 32-bit word offset in the workgroup's LDS group region of the
  // query's touched bitmap (batch_touch_words words), followed by its
  // accumulators for an LDS sink
  int acc;
  // dense group index (GPUProcessing.h:142) as hbase + sum_d field_d * uniq[d]
  // with field_d = (packed entry >> shift[d]) & (2^bits[d] - 1) and hbase =
  // sum over grouped dimensions of (attribute base - min_val) * unique_val;
  // bits[d] = uniq[d] = 0 for a dimension without group value (identical in
  // 32-bit wrap-around).  field_d < 2^16 and uniq[d] < 2^24, so the products
  // are 24-bit multiplies.
  int hbase;
  unsigned char expr;          // index into BatchDesc::expr
  unsigned char flags;         // BatchQueryFlags
  unsigned char sink;          // BatchSink
  unsigned char pad;
  unsigned char shift[4];
  unsigned char bits[4];
  int uniq[4];
  signed char attr[4];         // per dimension slot: -1 no group value, else its attribute
  int gtouch;                  // word offset of its touched bitmap in BatchDesc::touch_arena
  int unused[2];
};
static_assert(sizeof(BatchQuery) == 64, "BatchQuery is staged as 4 x 16 bytes");

// Group keys of a query whose keys decode from the group index (a mixed
// radix: digit d = value - minv[d] in [0, krange[d]) for uniq[d] != 0, key
// word 0 elsewhere), written after the scan for every touched group.
struct BatchKeys {
  int minv[4];
  int krange[4];
};

// 32-bit words of a query's touched bitmap (an even count, so that 64-bit
// accumulators after it stay 8-byte aligned); scalar aggregates and queries
// that store their keys in the pair loop keep none.
__host__ __device__ inline int batch_touch_words(int G, int flags) {
  return (flags & (kQueryAggr | kQueryKeysInLoop)) ? 0 : ((G + 63) >> 6) << 1;
}
// 32-bit words of a query's accumulators in LDS.
__host__ __device__ inline int batch_acc_words(int sink, int G) {
  return sink == kSinkLds64 ? 2 * G : sink == kSinkLds32 ? G : 0;
}

struct BatchPred {             // fact BETWEEN predicate of one query
  int col, lo, hi, q;
};

struct BatchExpr {
  int kind, a1, a2, pad;       // a2 < 0: absent (0)
};

struct BatchAttr {             // one packed attribute of a dimension
  const int* col;              // segment table of the dimension column
  int base;                    // smallest value
  int off, bits;               // position in the packed entry
  int lut;                     // offset of its mask table (one entry per value) or -1
  int span;                    // values base .. base + span - 1
};

struct BatchCmp {              // range filter of one query, evaluated per row:
  int src;                     // -1 = the key, else attribute index
  int lo, hi, q;               // src value in [lo, hi]
};

struct BatchDim {
  const int* key;              // segment table of the dimension key column
  void* pk;                    // packed table, len entries of pw bytes
  unsigned long long uses;     // queries that probe this dimension
  int used, len, minkey, nrows;
  int pw;                      // bytes per packed entry: 1, 2 or 4
  int nattr, ncmp, pad;
  // interval table of the range filters on one source (src as in BatchCmp)
  // when many queries filter it: ivl_n sorted breakpoints at lut[ivl_off ..]
  // (low 32 bits), then ivl_n + 1 query masks, one per interval between them
  int ivl_src, ivl_n, ivl_off, pad2;
  BatchAttr attr[kBatchMaxAttr];
  BatchCmp cmp[kBatchMaxCmp];
};

struct BatchDesc {
  int nq, ncols, npred, nexpr;
  unsigned long long all;      // queries of the batch (dead ones cleared)
  // row range of this launch: positions [0, num_tuples) of the segment group
  // starting at seg_group[seg_begin] (a chunk of a circular scan, or all)
  int seg_begin, num_tuples;
  int rebuild;                 // record tables must be rebuilt (query set changed)
  int pad0;
  unsigned long long reset;    // queries whose result rows are cleared first
  int lds_words;               // 32-bit words of the LDS group region
  int touch_words;             // words of the touch arena
  long long res_ints;          // result arena ints
  int* res_arena;
  unsigned* touch_arena;       // per query: a touched bitmap of its groups (gtouch)
  const int* col[kBatchMaxCols];
  int pred_begin[kBatchMaxCols + 1];
  BatchPred pred[kBatchMaxPred];
  BatchExpr expr[kBatchMaxExpr];
  BatchDim dim[4];
  BatchQuery q[kBatchMaxQ];
  BatchKeys keys[kBatchMaxQ];
  int nlut, pad1;
  unsigned long long lut[kBatchMaxLut];  // attribute mask tables
};

struct BatchLaunch {           // kernel arguments, fixed for a captured graph
  const BatchDesc* desc;
  const int* cache;
  const short* seg_group;
  int max_tuples;              // largest num_tuples of any launch (sizes the grid)
  unsigned* status;            // reserved status words (cleared by the fill)
};

// Hot part of the descriptor, staged in LDS once per workgroup so the tile
// loop reads it with broadcast LDS loads instead of dependent global loads.
struct BatchHotDim {
  const void* pk;
  unsigned long long uses;
  int len, minkey, pw, nattr, ncmp, pad;
  int ivl_src, ivl_n, ivl_off, pad2;
  int off[kBatchMaxAttr], bits[kBatchMaxAttr], lut[kBatchMaxAttr], base[kBatchMaxAttr];
};

struct BatchHot {
  const int* col[kBatchMaxCols];
  BatchHotDim dim[4];
  int ncols, colmask, usedmask, nexpr, npred, nq, nlut, pad;
  int pred_begin[kBatchMaxCols + 1];
  int4 expr[kBatchMaxExpr];
  int4 pred[kBatchMaxPred];    // lo, hi - lo (unsigned), query, column
  int4 cmp[4][kBatchMaxCmp];   // per dimension: src, lo, hi - lo (unsigned), query
};

struct BatchConfig {
  int block;                   // 256, 512, 1024
  int vpt;                     // 1, 2 (items per thread = 4*vpt)
  int lds_bytes;               // LDS per workgroup: staged descriptor, query
                               // descriptors and LDS group arrays
};

// Bytes of LDS a query's group array needs.
inline size_t batch_group_lds(int total_val) { return (size_t)total_val * 8; }

// Host launchers (batch.hip).  Every launch reads the batch from L.desc.
// rec / rec_len: packed tables and their sizes in 16-byte units
hipError_t batch_fill(const BatchLaunch& L, const uint4* const* rec, const int* rec_len, hipStream_t s);
hipError_t batch_records(const BatchLaunch& L, int d, int nrows, hipStream_t s);
// Writes the decoded keys of every touched group (after the scan).
hipError_t batch_keys(const BatchLaunch& L, hipStream_t s);
// ncols: fact columns the batch reads (selects the register layout).
hipError_t batch_scan(const BatchLaunch& L, const BatchConfig& c, int ncols, hipStream_t s);
bool batch_config_ok(const BatchConfig& c);
// Allow c.lds_bytes of dynamic LDS for the scan kernel (call before capture).
hipError_t batch_prepare(const BatchConfig& c, int ncols);
int batch_resident(const BatchConfig& c, int ncols);

}  // namespace mordred
