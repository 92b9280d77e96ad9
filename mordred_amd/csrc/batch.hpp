// Shared-scan execution of a batch of star queries (up to 64) on gfx950.
//
// One persistent kernel scans the fact columns once for the whole batch.  Per
// fact row it evaluates every query's fact predicates as bits of a 64-bit
// "alive" word, probes each joined dimension ONCE for all queries, and
// accumulates every (row, query) pair that survives.
//
// Per used dimension the batch builds tables indexed by the perfect-hash slot
// HASH = (key-min) % len of the reference's hash table (crystal/join.cuh:3),
// in one of two forms chosen by the table's size (every probe is a random
// gather, so what matters is how much of the table stays in L2):
//   * mask form (small dimensions): a query-mask table, 64 bits per slot, bit
//     q set iff a dimension row occupies the slot and passes query q's
//     dimension filter (or q does not filter this dimension), 0 for an empty
//     slot -- the union of the reference builds of all the batch's queries
//     (build_GPU2, GPUProcessing.h:927-997), evaluated once per dimension
//     row -- plus, when a query groups on the dimension, a table of the
//     row's group attributes packed at fixed bit positions;
//   * attribute form (large dimensions): one table of 1, 2 or 4 bytes per
//     slot holding every attribute the batch filters or groups on, each as
//     value - base + 1 at a fixed bit position, so that an empty slot is 0.
//     A filter on a narrow attribute reads a mask table in LDS (one 64-bit
//     query mask per value); a filter on the key is evaluated on the fact
//     row's key (the slot's row has that key) through the key column's
//     interval table.
// A fact row survives for query q iff every table q joins holds a row in the
// probed slot that passes q's filter, and its group value for table d is the
// attribute q groups on (the value word of q's own hash table), so every
// query's result equals its plain run bit for bit.  The host checks that keys
// cannot collide (dimension and fact key ranges within dim_len).
//
// The constraints on one fact column -- fact BETWEEN predicates, and key
// filters of attribute-form dimensions on their fact key -- become an interval
// table when there are enough of them: their ends split the column's value
// range into intervals, each with the mask of the queries that accept it; a
// row finds its interval from a bucket table (rank at the bucket start) plus
// at most kmax compares with the breakpoints inside the bucket.
//
// Everything that changes between batches lives in a device BatchDesc, so
// one captured hipGraph (fill, table builds, scan, keys) serves every batch
// of the same dataset and launch configuration.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mordred {

constexpr int kBatchMaxQ = 64;
constexpr int kBatchMaxCols = 12;   // fact columns; slots 0..3 are the dimension keys
constexpr int kBatchMaxPred = 2 * kBatchMaxQ;
constexpr int kBatchMaxExpr = 4;
constexpr int kBatchMaxAttr = 3;    // group attributes packed per dimension
constexpr int kBatchMaxFCol = 4;    // filtered dimension columns other than the key
constexpr int kBatchMaxFilt = 64;   // dimension filters per dimension (one per query)
constexpr int kBatchMaxIvl = 8;     // fact columns with an interval table
constexpr int kBatchIvlMin = 4;     // predicates on one fact column that get an interval table
constexpr int kBatchMaxBuckets = 512;
constexpr int kBatchMaxTab = 8192;  // 32-bit words of interval and mask tables (staged in LDS)
constexpr int kBatchLutSpan = 2048; // widest attribute given a mask table per value
constexpr long long kBatchMaskFormBytes = 1 << 20;  // dimensions whose mask table stays this small
// Dimension table forms: mask (per slot the 64-bit mask of the queries whose
// filters the row passes), attribute (per slot the packed attributes, pw
// bytes), dictionary (per slot a cb-bit code of the row's attribute tuple;
// an LDS dictionary maps a code to its packed attributes and to the mask of
// the queries whose attribute filters the tuple passes; code 0 = empty slot).
enum BatchDimForm { kDimMask = 0, kDimAttr = 1, kDimDict = 2 };
constexpr int kBatchDictMax = 1024;  // distinct attribute tuples of a dictionary (16 LDS bytes each)

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
// 16-byte LDS reads).  LDS offsets count from the start of the scan kernel's
// dynamic LDS, so that the pair loop's addresses are the offset plus an
// index.
struct BatchQuery {            // 64 bytes; staged in LDS by the scan kernel
  long long res;               // device address of the query's 6*total_val ints
  int total_val;
  int acc;                     // byte offset of the accumulators (LDS sinks, 8-byte aligned)
  // dense group index (GPUProcessing.h:142) as hbase + sum_d field_d * uniq[d]
  // with field_d = (attribute word >> shift[d]) & (2^bits[d] - 1) and hbase =
  // sum over joined dimensions of (group value base - min_val) * unique_val
  // (base 0 for a dimension without group value); bits[d] = 0 when the
  // dimension gives no group value.  field_d < 2^16 and uniq[d] < 2^24, so the
  // products are 24-bit multiplies (identical in 32-bit wrap-around).
  int hbase;
  unsigned char expr;          // index into BatchDesc::expr
  unsigned char sink;          // BatchSink
  // 32-bit word offset of the touched bitmap (batch_touch_words words), or
  // kBatchNoTouch for a query that keeps none
  unsigned short touch;
  unsigned char shift[4];
  unsigned char bits[4];
  int uniq[4];
  signed char attr[4];         // per dimension slot: -1 no group value, else its attribute
  int gtouch;                  // word offset of its touched bitmap in BatchDesc::touch_arena
  int flags;                   // BatchQueryFlags
  // the next query of the batch that is identical to this one and takes this
  // one's LDS sums, touched groups and keys instead of scanning (-1: none)
  int alias;
};
constexpr unsigned short kBatchNoTouch = 0xFFFF;
static_assert(sizeof(BatchQuery) == 64, "BatchQuery is staged as 4 x 16 bytes");
static_assert(offsetof(BatchQuery, alias) == 60, "the scan reads alias as the last word of chunk 3");

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

// The scan lists the (row, query) pairs of one row slot per wave before it
// aggregates them, so that the lanes of a wave share the pairs evenly: per
// wave a counter word and kBatchPairCap 16-bit entries (query << 8 | source
// lane << 2) at the end of the dynamic LDS region.  A slot with more pairs is
// listed in rounds.  The query descriptors are staged at its start, four
// 16-byte chunks per query at a stride of kBatchMaxQ chunks.
constexpr int kBatchPairCap = 256;
__host__ __device__ inline int batch_pair_bytes(int block) {
  return (block / 64) * (4 + 2 * kBatchPairCap);  // a multiple of 16
}

struct BatchPred {             // fact BETWEEN predicate of one query
  int col, lo, hi, q;
};

struct BatchExpr {
  int kind, a1, a2, pad;       // a2 < 0: absent (0)
};

struct BatchAttr {             // one packed attribute of a dimension
  const int* col;              // segment table of the dimension column
  int base;                    // smallest value (attribute form: stored as value - base + 1)
  int off, bits;               // position in the packed word
  int lut;                     // attribute form, filtered: word offset of its mask table in tab, else -1
  int pad;
};

struct BatchFilt {             // dimension filter of one query (build_GPU2's modes)
  int src;                     // -1: the key column, else an index into BatchDim::fcol
  int mode;                    // 0 EQ c1, 1 BETWEEN c1 c2, 2 c1 OR c2
  int c1, c2;
  int q;
  int pad[3];
};

struct BatchDim {
  const int* key;              // segment table of the dimension key column
  unsigned long long* mk;      // query-mask table, len entries of mstride u64 words
  void* pk;                    // attribute table (attribute form, len entries of pw bytes;
                               // dictionary form, len + 1 codes of cb bits), or nullptr
  unsigned long long uses;     // queries that probe this dimension
  unsigned long long groupers; // queries that take a group value from it
  int used, len, minkey, nrows;
  int pw;                      // bytes per attribute entry: 1, 2 or 4
  int nattr, nfcol, nfilt;
  int form;                    // BatchDimForm
  // mask form: u64 words per slot of mk; 2 when a query groups on the
  // dimension (the mask, then the packed group attributes in the next word's
  // low half: one 16-byte gather reads both)
  int mstride;
  int cb, dict_off;            // dictionary form: bits per code; LDS table word (a multiple
                               // of 4) of the dictionary: per code its mask, attributes, 0
  long long pk_words;          // 16-byte words of the attribute table (cleared by the fill)
  BatchAttr attr[kBatchMaxAttr];
  const int* fcol[kBatchMaxFCol];
  BatchFilt filt[kBatchMaxFilt];
};

// Interval table of the predicates on one fact column, in the LDS table
// region (32-bit words from BatchDesc::tab): masks (n + 1 u64 at an even
// offset), bucket ranks (nb u16, two per word), breakpoints (n + kmax ints,
// padded with INT_MAX).  Bucket t covers values base + [t << shift,
// (t+1) << shift); its rank is the number of breakpoints below its start.
struct BatchIvl {
  int col;                     // fact column slot
  int base, shift, nb;
  int n, kmax;
  int mask_off, rank_off, bp_off;
  int pad[3];
};

struct BatchDesc {
  int nq, ncols, npred, nexpr;
  unsigned long long all;      // queries of the batch (dead ones cleared)
  // row range of this launch: positions [0, num_tuples) of the segment group
  // starting at seg_group[seg_begin] (a chunk of a circular scan, or all)
  int seg_begin, num_tuples;
  int rebuild;                 // dimension tables must be rebuilt (query set changed)
  int nivl;
  unsigned long long reset;    // queries whose result rows are cleared first
  unsigned long long aliases;  // queries that take an identical query's sums (not in `all`)
  int lds_words;               // 32-bit words of the LDS group region
  int touch_words;             // words of the touch arena
  long long res_ints;          // result arena ints
  int* res_arena;
  unsigned* touch_arena;       // per query: a touched bitmap of its groups (gtouch)
  const int* col[kBatchMaxCols];
  int pred_begin[kBatchMaxCols + 1];  // generic predicates (not in an interval table), by column
  BatchPred pred[kBatchMaxPred];
  BatchExpr expr[kBatchMaxExpr];
  BatchIvl ivl[kBatchMaxIvl];
  BatchDim dim[4];
  BatchQuery q[kBatchMaxQ];
  BatchKeys keys[kBatchMaxQ];
  int ntab;
  int lds_dyn;                 // dynamic LDS bytes of the scan launch (its pair lists end there)
  // every dimension's attribute word fits in 16 bits: the scan packs dimensions
  // 1 and 3 into the high halves of 0 and 2 (BatchQuery::shift of 1 and 3
  // count from bit 16)
  int attr16;
  // every query's uniq[d] is below 2^16: the scan stages each descriptor with
  // the pair step's operands in its first two 16-byte chunks (uniq as 16-bit
  // halves) and the result address in the third
  int uniq16;
  unsigned tab[kBatchMaxTab];  // interval tables and attribute mask tables
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
  const unsigned long long* mk;
  const void* pk;
  unsigned long long uses, groupers;
  int len, minkey, pw, form;
  int nattr, cb;
  int base[kBatchMaxAttr], dict;  // group attribute bases (keys stored in the pair loop)
  int off[kBatchMaxAttr], bits[kBatchMaxAttr], lut[kBatchMaxAttr], mstride, pad3[2];
};

struct BatchHot {
  const int* col[kBatchMaxCols];
  BatchHotDim dim[4];
  int ncols, colmask, usedmask, nexpr, npred, nq, ntab, nivl, attr16, pad[3];
  int pred_begin[kBatchMaxCols + 1];
  int4 expr[kBatchMaxExpr];
  int4 pred[kBatchMaxPred];    // lo, hi - lo (unsigned), query, column
  BatchIvl ivl[kBatchMaxIvl];
};

static_assert(sizeof(BatchHot) % 16 == 0, "the scan's dynamic LDS (and its pair lists) stay 16-byte aligned");

struct BatchConfig {
  int block;                   // 256, 512, 1024
  int vpt;                     // 1, 2 (items per thread = 4*vpt)
  int lds_bytes;               // LDS per workgroup: staged descriptor, query
                               // descriptors, interval tables and group region
};

// Host launchers (batch.hip).  Every launch reads the batch from L.desc.
hipError_t batch_fill(const BatchLaunch& L, hipStream_t s);
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
