// gfx950 query kernels and their C ABI (include/mordred_hip.h).
//
// One generic tile pipeline replaces the 18 fused kernel templates of
// src/gpudb/GPUProcessing.h.  A pipeline is
//     row source  ->  filters  ->  perfect-hash probes  ->  sink
// with
//   row source: SEG   segment mode, rows of a tile are one contiguous run of
//                     one cached segment  (the *_GPU2 kernels)
//               OFF   offset mode, rows come from a row-id list and every
//                     column value is a segment-indirected gather (*_GPU3)
//               PLAIN plain column pointers (the on-demand kernels)
//   sink:       GROUPBY_LDS  dense int64 group array kept in LDS per
//                            workgroup, flushed once with global atomics
//               GROUPBY_GLOBAL  one global 64-bit atomic per selected row
//                            (group arrays too large for LDS)
//               AGGR         int64 sum of a1*a2, one atomic per workgroup
//               COMPACT      selected row ids (+ dimension rows) written to
//                            offset lists, one reservation atomic per tile
// Filters, probed tables and outputs are selected at run time by NULL-ness of
// the argument pointers exactly as in the reference kernels; every such test
// is workgroup-uniform, so it costs one scalar branch.
//
// Workgroups are persistent: the grid is sized to the number of workgroups the
// device keeps resident, and each workgroup walks tiles blockIdx.x,
// blockIdx.x + gridDim.x, ...  A tile is BLOCK*4*VPT rows and never straddles
// a 2^20-row segment.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <vector>

#define MORDRED_DEFINE_PREP_KERNELS
#include "pipeline.hpp"
#include "stream_pack.hpp"

namespace mordred {

// ----------------------------------------------------------------- host side

struct OccKey {
  const void* f;
  size_t lds;
  int dev;
  bool operator==(const OccKey& o) const { return f == o.f && lds == o.lds && dev == o.dev; }
};
struct OccHash {
  size_t operator()(const OccKey& k) const {
    return std::hash<const void*>()(k.f) ^ (k.lds * 0x9E3779B97F4A7C15ull) ^ (size_t)k.dev;
  }
};

// Number of workgroups the device keeps resident for kernel f (cached).
int resident_groups(const void* f, int block, size_t lds) {
  static std::mutex mu;
  static std::unordered_map<OccKey, int, OccHash> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  OccKey key{f, lds, dev};
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int cus = 0, occ = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, block, lds) != hipSuccess || occ <= 0)
    occ = 1;
  int g = cus * occ;
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = g;
  return g;
}

// Stream-ordered scratch for the compact probe tables: one memory pool per
// device, created on first use and kept (freed blocks stay cached for reuse).
static hipMemPool_t device_pool(int dev) {
  static std::mutex mu;
  static std::unordered_map<int, hipMemPool_t> pools;
  std::lock_guard<std::mutex> g(mu);
  auto it = pools.find(dev);
  if (it != pools.end()) return it->second;
  hipMemPoolProps props{};
  props.allocType = hipMemAllocationTypePinned;
  props.handleTypes = hipMemHandleTypeNone;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = dev;
  hipMemPool_t pool = nullptr;
  if (hipMemPoolCreate(&pool, &props) != hipSuccess) pool = nullptr;
  if (pool) {
    uint64_t keep = ~0ull;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  pools[dev] = pool;
  return pool;
}

static bool probed(const Pipe& P, int k) { return (P.pmask >> k & 1) && P.key[k] && P.ht[k]; }

// Compile-time shape of a segment-mode fused group-by or aggregate (see
// pipeline_kernel), or kShapeGeneric.
static int shape_of(const Pipe& P, int mode, int sink) {
  if (mode != SEG || (sink != GROUPBY_LDS && sink != AGGR)) return kShapeGeneric;
  if (P.fkind != F_NONE && P.fkind != F_BETWEEN) return kShapeGeneric;
  int s = 0;
  for (int k = 0; k < 4; k++)
    if (probed(P, k)) s |= 1 << k;
  if (P.fkind == F_BETWEEN) {
    if (P.fcol[0]) s |= 0x10;
    if (P.fcol[1]) s |= 0x20;
  }
  return s;
}

struct ProbePlan {
  int compact;    // bit k: build a compact table for table k
  int stage;      // bit k: table k may be staged in LDS
  int lds_words;  // LDS (32-bit words) reserved for staged tables
};

static int sink_block(int sink) {
  return sink == GROUPBY_LDS ? kLdsBlock : sink == AGGR ? kAggrBlock : kBlock;
}

// Dynamic LDS a workgroup of `sink` may use.
static size_t lds_cap(int sink) { return sink == GROUPBY_LDS ? kLdsBudgetOneGroup : kLdsBudget; }

// Which tables get a compact probe table, and which may be staged in LDS.
// A table is compacted when the fact scan is long next to it (>= 16 rows per
// slot, so the two pre-pass launches are paid back) and either it is probed by
// every row (the first probe with no filter before it), or it is too large to
// stay in L2, or (>= 256 rows per slot) its packed form may fit in the LDS the
// group array and the survivor rings leave.  A compacted table whose 1-bit form
// fits may be staged; the kernel stages tables in probe order while their
// actual packed words fit in lds_words.
static ProbePlan plan_probes(const Pipe& P, int mode, int sink) {
  ProbePlan plan{0, 0, 0};
  const bool filtered = P.fkind != F_NONE && (P.fcol[0] || P.fcol[1]);
  size_t fixed = group_lds_bytes(sink, P.total_val);
  const int shape = shape_of(P, mode, sink);
  if (shape_compacts(mode, sink, shape)) fixed += ring_lds_bytes(sink_block(sink), shape);
  const size_t avail = lds_cap(sink) > fixed ? lds_cap(sink) - fixed : 0;
  size_t want = 0;
  int first = -1;
  for (int k = 0; k < 4; k++) {
    if (!probed(P, k)) continue;
    const bool is_first = first < 0;
    if (is_first) first = k;
    const int len = P.len[k];
    if (len <= 0 || (long long)P.num_tuples < 16LL * len) continue;
    const bool stageable = staged_words(len, 1) * 4 <= avail;
    const bool lds_worth = (long long)P.num_tuples >= 256LL * len && staged_words(len, 4) * 4 <= avail;
    if ((is_first && !filtered) || len >= kCompactMinLen || lds_worth) {
      plan.compact |= 1 << k;
      if (stageable && (lds_worth || (is_first && !filtered))) {
        plan.stage |= 1 << k;
        // one workgroup per CU: take all that is left; otherwise reserve for
        // entries of up to 4 bits, so that LDS does not cut occupancy
        want += staged_words(len, sink == GROUPBY_LDS ? 16 : 4) * 4;
      }
    }
  }
  plan.lds_words = (int)(std::min(avail, want) / 4);
  return plan;
}

template <int MODE>
static int run_pipe(const Pipe& P, int sink, hipStream_t stream) {
  const int shape = shape_of(P, MODE, sink);
  if (shape != kShapeGeneric) {
    const int rc = sink == AGGR ? launch_aggr_spec(P, shape, stream) : launch_groupby_lds_spec(P, shape, stream);
    if (rc != -1) return rc;
  }
  switch (sink) {
    // offset-mode group-by gathers every column by row: 4 rows per thread
    // keep its registers within the 1024-thread budget (8 spill to scratch)
    case GROUPBY_LDS:
      if constexpr (MODE == OFF) return launch_pipe<kLdsBlock, 1, MODE, GROUPBY_LDS>(P, stream);
      else return launch_pipe<kLdsBlock, kVpt, MODE, GROUPBY_LDS>(P, stream);
    case GROUPBY_GLOBAL: return launch_pipe<kBlock, kVpt, MODE, GROUPBY_GLOBAL>(P, stream);
    case AGGR: return launch_pipe<kAggrBlock, kAggrVpt, MODE, AGGR>(P, stream);
    default: return launch_pipe<kBlock, kVpt, MODE, COMPACT>(P, stream);
  }
}

// Bytes of one compact probe table of `len` slots: a 256-byte header, the
// stats partials, the packed words at the widest width, and (large tables)
// the existence bitmap.
static size_t compact_table_bytes(int len) {
  size_t b = 256 + kStatsGroups * sizeof(int4) + ((compact_words_bytes(len) + 255) & ~(size_t)255);
  if (len >= kCompactMinLen) b += (compact_bits_bytes(len) + 255) & ~(size_t)255;
  return b;
}

// Points slot nt of `prep` at a compact table of `len` slots laid out at t.
static void prep_table(Prep& prep, int nt, const int* ht, int len, char* t) {
  prep.ht[nt] = ht;
  prep.len[nt] = len;
  prep.hdr[nt] = reinterpret_cast<CompactHdr*>(t);
  prep.partial[nt] = reinterpret_cast<int4*>(t + 256);
  char* w = t + 256 + kStatsGroups * sizeof(int4);
  prep.words[nt] = reinterpret_cast<uint32_t*>(w);
  prep.bits[nt] = len >= kCompactMinLen
                      ? reinterpret_cast<uint32_t*>(w + ((compact_words_bytes(len) + 255) & ~(size_t)255))
                      : nullptr;
}

// The two pre-pass launches over the nt tables of `prep`.
static hipError_t launch_prep_kernels(const Prep& prep, int nt, hipStream_t stream) {
  prep_stats_kernel<<<dim3(kStatsGroups, nt), kPrepBlock, 0, stream>>>(prep);
  int maxlen = 0;
  for (int t = 0; t < nt; t++) maxlen = std::max(maxlen, prep.len[t]);
  const int pack_groups = std::max(1, std::min(kPackGroups, (maxlen + kPrepBlock - 1) / kPrepBlock));
  prep_pack_kernel<<<dim3(pack_groups, nt), kPrepBlock, 0, stream>>>(prep);
  return hipGetLastError();
}

// Byte layout of the compact probe tables of the tables in `mask`, one after
// the other in one scratch block.
struct CompactLayout {
  size_t off[4];
  size_t bytes;
};

static CompactLayout compact_layout(const Pipe& P, int mask) {
  CompactLayout L{};
  for (int k = 0; k < 4; k++) {
    if (!(mask >> k & 1)) continue;
    L.off[k] = L.bytes;
    L.bytes += compact_table_bytes(P.len[k]);
  }
  return L;
}

// Launch the two pre-pass kernels that derive the compact tables of the
// tables in `mask` into `base` (laid out by L) and point the pipe at them.
// Returns the launch status; on failure those tables are probed in place.
static hipError_t launch_prep(Pipe& P, int mask, const CompactLayout& L, char* base, int sink, hipStream_t stream) {
  Prep prep{};
  prep.row_probe = sink == COMPACT;
  int nt = 0;
  for (int k = 0; k < 4; k++) {
    if (!(mask >> k & 1)) continue;
    prep_table(prep, nt, P.ht[k], P.len[k], base + L.off[k]);
    P.chdr[k] = prep.hdr[nt];
    P.cwords[k] = prep.words[nt];
    P.cbits[k] = prep.bits[nt];
    nt++;
  }
  if (nt == 0) return hipSuccess;
  const hipError_t e = launch_prep_kernels(prep, nt, stream);
  if (e != hipSuccess)
    for (int k = 0; k < 4; k++)
      if (mask >> k & 1) { P.chdr[k] = nullptr; P.cwords[k] = nullptr; P.cbits[k] = nullptr; }
  return e;
}

// Build the compact probe tables of the tables in `mask` into pool scratch
// (stream-ordered).  Returns the scratch block (to free after the fact
// kernel) or nullptr, in which case the fact kernel probes those hash tables
// in place.
static void* build_compact(Pipe& P, int mask, int sink, hipStream_t stream) {
  if (!mask) return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  hipMemPool_t pool = device_pool(dev);
  if (!pool) return nullptr;
  const CompactLayout L = compact_layout(P, mask);
  void* scratch = nullptr;
  if (hipMallocFromPoolAsync(&scratch, L.bytes, pool, stream) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (launch_prep(P, mask, L, static_cast<char*>(scratch), sink, stream) != hipSuccess) {
    (void)hipFreeAsync(scratch, stream);
    return nullptr;
  }
  return scratch;
}

// ------------------------------------------------- registered probe tables
// A caller that probes one hash table in many fused calls (segment groups,
// streamed chunks, pipeline stages of one query) registers it.  The compact
// forms that those calls would each derive (one for value probes, one for
// row probes) are then derived once, by the first call that compacts the
// table, into memory allocated at registration, and reused by later calls on
// any stream (they wait for the derivation's event) until a build into the
// table (any mordred_build_* entry point) marks them stale.  Whether a call
// compacts a table at all stays the per-call plan's decision.
struct RegTable {
  int dev;
  const int* ht;
  int len;
  char* mem;             // [value-probe form | row-probe form]
  size_t form_bytes;
  bool fresh[2];         // derived since the table's last build
  hipEvent_t ready[2];   // recorded after each form's last derivation
};
static std::mutex reg_mu;
static std::vector<RegTable> reg_tables;

static RegTable* reg_find(int dev, const int* ht) {
  for (RegTable& t : reg_tables)
    if (t.dev == dev && t.ht == ht) return &t;
  return nullptr;
}

static void reg_free(RegTable& t) {
  for (int f = 0; f < 2; f++)
    if (t.ready[f]) (void)hipEventDestroy(t.ready[f]);
  if (t.mem) (void)hipFree(t.mem);
}

// A build into `ht` makes its registered forms stale.
static void reg_mark_stale(const int* ht) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> g(reg_mu);
  if (RegTable* t = reg_find(dev, ht)) t->fresh[0] = t->fresh[1] = false;
}

// Registered tables among `mask`: point the pipe at their forms, deriving a
// stale form on `stream` first.  `done` gets the tables served this way.
// Allocates nothing.
static hipError_t use_registered(Pipe& P, int mask, int row_probe, hipStream_t stream, int& done) {
  done = 0;
  if (!mask) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(reg_mu);
  if (reg_tables.empty()) return hipSuccess;
  for (int k = 0; k < 4; k++) {
    if (!(mask >> k & 1)) continue;
    RegTable* t = reg_find(dev, P.ht[k]);
    if (!t || t->len != P.len[k]) continue;
    Prep prep{};
    prep.row_probe = row_probe;
    prep_table(prep, 0, t->ht, t->len, t->mem + (size_t)row_probe * t->form_bytes);
    if (!t->fresh[row_probe]) {
      // derive under the lock, so that a call on another stream that finds
      // the form fresh waits for an event already recorded after it
      if ((e = launch_prep_kernels(prep, 1, stream)) != hipSuccess) return e;
      if ((e = hipEventRecord(t->ready[row_probe], stream)) != hipSuccess) return e;
      t->fresh[row_probe] = true;
    } else if ((e = hipStreamWaitEvent(stream, t->ready[row_probe], 0)) != hipSuccess) {
      return e;
    }
    P.chdr[k] = prep.hdr[0];
    P.cwords[k] = prep.words[0];
    P.cbits[k] = prep.bits[0];
    done |= 1 << k;
  }
  return hipSuccess;
}

// ------------------------------------------------- fixed scratch (no-alloc)
// Device memory of the no-alloc entry point: one array per GPU, of a fixed
// size, held until the process ends.  In host code `scratch` is a host-side
// stand-in, not a device address: host code uses only the per-device address
// that hipGetSymbolAddress returns (taken once per device, with that device
// current).  Every call on a GPU shares that GPU's array; the pre-pass
// kernels rewrite every byte a call reads (headers, partials, packed words,
// bitmaps), so nothing needs clearing between calls, and calls that overlap
// on one GPU race.
static __device__ __align__(256) unsigned char scratch[MORDRED_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

// Address of the current device's scratch array, or nullptr (err set).
static char* device_scratch(hipError_t& err) {
  static std::mutex mu;
  static void* ptrs[kMaxDevices] = {};
  int dev = 0;
  if ((err = hipGetDevice(&dev)) != hipSuccess) return nullptr;
  if (dev < 0 || dev >= kMaxDevices) {
    err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> g(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    if ((err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return nullptr;
    ptrs[dev] = p;
  }
  err = hipSuccess;
  return static_cast<char*>(ptrs[dev]);
}

// The group index sum_k (v_k - minv_k) * uniq_k is a mixed-radix number when
// the non-zero strides, in increasing order, each divide the next and the
// largest divides total_val: digit k then ranges over [0, next stride /
// uniq_k), and keys whose digits lie in range (the kernel checks each row)
// decode exactly from the index.
static void set_key_decode(Pipe& P) {
  P.kdecode = 0;
  int idx[4], m = 0;
  for (int k = 0; k < 4; k++) {
    P.krange[k] = 0;
    if (P.uniq[k] < 0) return;
    if (P.uniq[k] > 0) idx[m++] = k;
  }
  if (P.total_val <= 0) return;
  std::stable_sort(idx, idx + m, [&](int a, int b) { return P.uniq[a] < P.uniq[b]; });
  for (int i = 0; i < m; i++) {
    const int next = i + 1 < m ? P.uniq[idx[i + 1]] : P.total_val;
    if (next % P.uniq[idx[i]] != 0) return;
    P.krange[idx[i]] = next / P.uniq[idx[i]];
  }
  P.kdecode = 1;
}

// sink < 0: group-by, placed in LDS when the dense array fits.
// fixed != nullptr: the compact probe tables go into that block of `cap`
// bytes (the no-alloc entry point); a plan that needs more is refused with
// hipErrorOutOfMemory before anything is launched.  Otherwise they go into
// stream-ordered pool memory, or are skipped if none is available.
static int dispatch(Pipe P, int mode, int sink, hipStream_t stream, char* fixed = nullptr, size_t cap = 0) {
  if (P.num_tuples == 0) return hipSuccess;
  if (sink < 0) set_key_decode(P);
  if (sink < 0) {
    const int shape = shape_of(P, mode, GROUPBY_LDS);
    const size_t ring = shape_compacts(mode, GROUPBY_LDS, shape) ? ring_lds_bytes(kLdsBlock, shape) : 0;
    sink = group_lds_bytes(GROUPBY_LDS, P.total_val) + ring <= kLdsBudgetOneGroup ? GROUPBY_LDS : GROUPBY_GLOBAL;
  }
  const ProbePlan plan = plan_probes(P, mode, sink);
  void* pooled = nullptr;
  int reg = 0;
  if (fixed) {
    // the refusal is decided before anything is launched
    if (compact_layout(P, plan.compact).bytes > cap) return hipErrorOutOfMemory;
  }
  const hipError_t re = use_registered(P, plan.compact, sink == COMPACT, stream, reg);
  if (re != hipSuccess) return re;
  const int rest = plan.compact & ~reg;
  bool compacted = reg != 0;
  if (fixed) {
    if (rest) {
      const hipError_t e = launch_prep(P, rest, compact_layout(P, rest), fixed, sink, stream);
      if (e != hipSuccess) return e;
      compacted = true;
    }
  } else if (rest) {
    pooled = build_compact(P, rest, sink, stream);
    compacted = compacted || pooled != nullptr;
  }
  if (compacted && plan.stage) {
    P.lds_stage = plan.stage;
    P.lds_words = plan.lds_words;
  }
  int rc;
  if (mode == SEG) rc = run_pipe<SEG>(P, sink, stream);
  else if (mode == OFF) rc = run_pipe<OFF>(P, sink, stream);
  else rc = run_pipe<PLAIN>(P, sink, stream);
  if (pooled) {
    const hipError_t e = hipFreeAsync(pooled, stream);
    if (rc == hipSuccess) rc = e;
  }
  return rc;
}

static bool aligned16(const void* p) { return p == nullptr || ((uintptr_t)p & 15u) == 0; }

static Pipe make_pipe(const int* cache, int num_tuples) {
  Pipe P{};
  P.cache = cache;
  P.num_tuples = num_tuples;
  P.aligned = aligned16(cache);  // SEG mode streams from the cache
  return P;
}

static void put_filter_between(Pipe& P, const filterArgsGPU* f, int kind) {
  P.fkind = kind;
  if (!f) return;
  P.fcol[0] = f->filter_idx1;
  P.fcol[1] = f->filter_idx2;
  P.cmp[0] = f->compare1; P.cmp[1] = f->compare2; P.cmp[2] = f->compare3; P.cmp[3] = f->compare4;
  P.fmode[0] = f->mode1; P.fmode[1] = f->mode2;
}

static void put_probe(Pipe& P, const probeArgsGPU* p, int pmask) {
  P.pmask = pmask;
  if (!p) return;
  P.key[0] = p->key_idx1; P.key[1] = p->key_idx2; P.key[2] = p->key_idx3; P.key[3] = p->key_idx4;
  P.ht[0] = p->ht1; P.ht[1] = p->ht2; P.ht[2] = p->ht3; P.ht[3] = p->ht4;
  P.len[0] = p->dim_len1; P.len[1] = p->dim_len2; P.len[2] = p->dim_len3; P.len[3] = p->dim_len4;
  P.minkey[0] = p->min_key1; P.minkey[1] = p->min_key2; P.minkey[2] = p->min_key3; P.minkey[3] = p->min_key4;
}

static void put_groupby(Pipe& P, const groupbyArgsGPU* g) {
  if (!g) return;
  P.aggr[0] = g->aggr_idx1; P.aggr[1] = g->aggr_idx2;
  P.grp[0] = g->group_idx1; P.grp[1] = g->group_idx2; P.grp[2] = g->group_idx3; P.grp[3] = g->group_idx4;
  P.minv[0] = g->min_val1; P.minv[1] = g->min_val2; P.minv[2] = g->min_val3; P.minv[3] = g->min_val4;
  P.uniq[0] = g->unique_val1; P.uniq[1] = g->unique_val2; P.uniq[2] = g->unique_val3; P.uniq[3] = g->unique_val4;
  P.total_val = g->total_val;
}

static void put_in_offsets(Pipe& P, const offsetGPU* o) {
  if (!o) return;
  P.in_lo = o->lo_off;
  P.in_dim[0] = o->dim_off1; P.in_dim[1] = o->dim_off2; P.in_dim[2] = o->dim_off3; P.in_dim[3] = o->dim_off4;
  P.aligned = aligned16(o->lo_off) && aligned16(o->dim_off1) && aligned16(o->dim_off2) &&
              aligned16(o->dim_off3) && aligned16(o->dim_off4);
}

static void put_out_offsets(Pipe& P, const offsetGPU* o) {
  if (!o) return;
  P.out[0] = o->lo_off;
  P.out[1] = o->dim_off1; P.out[2] = o->dim_off2; P.out[3] = o->dim_off3; P.out[4] = o->dim_off4;
}

// Segment-mode addressing: a NULL segment group needs a segment-aligned start.
static bool seg_ok(const int* cache, int num_tuples, int start_offset, const short* sg) {
  if (num_tuples < 0) return false;
  if (num_tuples > 0 && !cache) return false;
  return sg != nullptr || (start_offset & kSegMask) == 0;
}

// Offset-mode group-by takes a group value from dim_offk when table k is not
// probed and group_idxk is set (GPUProcessing.h:202-206 asserts dim_offk).
static bool groupby_inputs_ok(const Pipe& P) {
  for (int k = 0; k < 4; k++)
    if (!(P.key[k] && P.ht[k]) && P.grp[k] && !P.in_dim[k]) return false;
  return true;
}

static constexpr int kGroupBySink = -1;

#ifndef MORDRED_BUILD_BLOCK
#define MORDRED_BUILD_BLOCK 256
#endif
#ifndef MORDRED_BUILD_VPT
#define MORDRED_BUILD_VPT 1
#endif
constexpr int kBuildBlock = MORDRED_BUILD_BLOCK;
constexpr int kBuildVpt = MORDRED_BUILD_VPT;

template <int MODE>
static int run_build(const Build& B, hipStream_t stream) {
  if (B.num_tuples == 0) return hipSuccess;
  reg_mark_stale(B.ht);
  constexpr int TILE = kBuildBlock * 4 * kBuildVpt;
  const void* f = reinterpret_cast<const void*>(&build_kernel<kBuildBlock, kBuildVpt, MODE>);
  const int ntiles = (B.num_tuples + TILE - 1) / TILE;
  const int grid = balanced_grid(ntiles, resident_groups(f, kBuildBlock, 0));
  build_kernel<kBuildBlock, kBuildVpt, MODE><<<grid, kBuildBlock, 0, stream>>>(B);
  return hipGetLastError();
}

// Bit-packed segments into their staging slots: one thread per value; a
// field spans at most two words (bits <= 24), the second always in bounds
// (each packed segment keeps a slack word).
__global__ __launch_bounds__(256) void unpack_kernel(UnpackBatch u) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= u.nseg || j >= u.rows[i]) return;
  const uint32_t* w = u.packed + (size_t)i * (size_t)packed_words(u.bits);
  const unsigned long long bit = (unsigned long long)j * (unsigned)u.bits;
  const unsigned long long pair = (unsigned long long)w[bit >> 5] | (unsigned long long)w[(bit >> 5) + 1] << 32;
  const uint32_t v = (uint32_t)(pair >> (bit & 31)) & ((1u << u.bits) - 1u);
  u.cache[((size_t)(unsigned)u.slot[i] << kPackSegLog) + j] = (int)((uint32_t)u.base + v);
}

hipError_t unpack_segments(const UnpackBatch& u, hipStream_t s) {
  if (u.nseg <= 0 || u.nseg > kPackMaxSegs || u.bits <= 0 || u.bits > kPackMaxBits) return hipErrorInvalidValue;
  int rmax = 0;
  for (int i = 0; i < u.nseg; i++) {
    if (u.rows[i] < 0 || u.rows[i] > (1 << kPackSegLog) || u.slot[i] < 0) return hipErrorInvalidValue;
    rmax = std::max(rmax, u.rows[i]);
  }
  if (rmax == 0) return hipSuccess;
  unpack_kernel<<<dim3((rmax + 255) / 256, u.nseg), 256, 0, s>>>(u);
  return hipGetLastError();
}

}  // namespace mordred

using namespace mordred;

#define INVALID_IF(c) \
  do {                \
    if (c) return hipErrorInvalidValue; \
  } while (0)

extern "C" {

const char* mordred_version(void) { return "mordred-hip 0.1 (gfx950)"; }

int mordred_probe_table_register(const int* hash_table, int dim_len) {
  INVALID_IF(!hash_table || dim_len <= 0);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(reg_mu);
  RegTable* t = reg_find(dev, hash_table);
  if (t && t->len == dim_len) return hipSuccess;
  if (t) {
    reg_free(*t);
    reg_tables.erase(reg_tables.begin() + (t - reg_tables.data()));
  }
  RegTable n{};
  n.dev = dev;
  n.ht = hash_table;
  n.len = dim_len;
  n.form_bytes = compact_table_bytes(dim_len);
  if ((e = hipMalloc(&n.mem, 2 * n.form_bytes)) != hipSuccess) return e;
  for (int f = 0; f < 2; f++) {
    if ((e = hipEventCreateWithFlags(&n.ready[f], hipEventDisableTiming)) != hipSuccess) {
      reg_free(n);
      return e;
    }
  }
  reg_tables.push_back(n);
  return hipSuccess;
}

int mordred_probe_table_unregister(const int* hash_table) {
  INVALID_IF(!hash_table);
  int dev = 0;
  const hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(reg_mu);
  RegTable* t = reg_find(dev, hash_table);
  if (!t) return hipSuccess;
  reg_free(*t);
  reg_tables.erase(reg_tables.begin() + (t - reg_tables.data()));
  return hipSuccess;
}

int mordred_probe_table_invalidate(const int* hash_table) {
  INVALID_IF(!hash_table);
  reg_mark_stale(hash_table);
  return hipSuccess;
}

int mordred_filter_probe_group_by_GPU2(const int* gpuCache, const filterArgsGPU* fargs,
    const probeArgsGPU* pargs, const groupbyArgsGPU* gargs, int num_tuples, int* res,
    int start_offset, const short* segment_group, mordred_stream_t stream) {
  INVALID_IF(!seg_ok(gpuCache, num_tuples, start_offset, segment_group) || !gargs || !res ||
             gargs->total_val <= 0);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.seg_group = segment_group; P.start_offset = start_offset; P.res = res;
  put_filter_between(P, fargs, F_BETWEEN);
  put_probe(P, pargs, 0xF);
  put_groupby(P, gargs);
  return dispatch(P, SEG, kGroupBySink, (hipStream_t)stream);
}

int mordred_filter_probe_group_by_GPU2_noalloc(const int* gpuCache, const filterArgsGPU* fargs,
    const probeArgsGPU* pargs, const groupbyArgsGPU* gargs, int num_tuples, int* res,
    int start_offset, const short* segment_group, mordred_stream_t stream) {
  INVALID_IF(!seg_ok(gpuCache, num_tuples, start_offset, segment_group) || !gargs || !res ||
             gargs->total_val <= 0);
  hipError_t err;
  char* fixed = device_scratch(err);
  if (!fixed) return err;
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.seg_group = segment_group; P.start_offset = start_offset; P.res = res;
  put_filter_between(P, fargs, F_BETWEEN);
  put_probe(P, pargs, 0xF);
  put_groupby(P, gargs);
  return dispatch(P, SEG, kGroupBySink, (hipStream_t)stream, fixed, MORDRED_NOALLOC_SCRATCH_BYTES);
}

int mordred_filter_probe_group_by_GPU3(const int* gpuCache, const offsetGPU* offset,
    const filterArgsGPU* fargs, const probeArgsGPU* pargs, const groupbyArgsGPU* gargs,
    int num_tuples, int* res, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || !offset || !offset->lo_off || !gargs || !res || gargs->total_val <= 0);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.res = res;
  put_in_offsets(P, offset);
  put_filter_between(P, fargs, F_BETWEEN);
  put_probe(P, pargs, 0xF);
  put_groupby(P, gargs);
  INVALID_IF(!groupby_inputs_ok(P));
  return dispatch(P, OFF, kGroupBySink, (hipStream_t)stream);
}

int mordred_probe_group_by_GPU2(const int* gpuCache, const probeArgsGPU* pargs,
    const groupbyArgsGPU* gargs, int num_tuples, int* res, int start_offset,
    const short* segment_group, mordred_stream_t stream) {
  return mordred_filter_probe_group_by_GPU2(gpuCache, nullptr, pargs, gargs, num_tuples, res,
                                            start_offset, segment_group, stream);
}

int mordred_probe_group_by_GPU3(const int* gpuCache, const offsetGPU* offset,
    const probeArgsGPU* pargs, const groupbyArgsGPU* gargs, int num_tuples, int* res,
    mordred_stream_t stream) {
  return mordred_filter_probe_group_by_GPU3(gpuCache, offset, nullptr, pargs, gargs, num_tuples,
                                            res, stream);
}

int mordred_groupByGPU(const int* gpuCache, const offsetGPU* offset, const groupbyArgsGPU* gargs,
    int num_tuples, int* res, mordred_stream_t stream) {
  return mordred_filter_probe_group_by_GPU3(gpuCache, offset, nullptr, nullptr, gargs, num_tuples,
                                            res, stream);
}

int mordred_filter_probe_GPU2(const int* gpuCache, const filterArgsGPU* fargs,
    const probeArgsGPU* pargs, const offsetGPU* out_off, int num_tuples, int* total,
    int start_offset, const short* segment_group, mordred_stream_t stream) {
  INVALID_IF(!seg_ok(gpuCache, num_tuples, start_offset, segment_group) || !total || !out_off ||
             !out_off->lo_off);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.seg_group = segment_group; P.start_offset = start_offset; P.total = total;
  put_filter_between(P, fargs, F_BETWEEN);
  put_probe(P, pargs, 0x8);
  put_out_offsets(P, out_off);
  return dispatch(P, SEG, COMPACT, (hipStream_t)stream);
}

int mordred_filter_probe_GPU3(const int* gpuCache, const offsetGPU* in_off,
    const filterArgsGPU* fargs, const probeArgsGPU* pargs, const offsetGPU* out_off,
    int num_tuples, int* total, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || !in_off || !in_off->lo_off || !total || !out_off || !out_off->lo_off);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.total = total;
  put_in_offsets(P, in_off);
  put_filter_between(P, fargs, F_BETWEEN);
  put_probe(P, pargs, 0x8);
  put_out_offsets(P, out_off);
  return dispatch(P, OFF, COMPACT, (hipStream_t)stream);
}

int mordred_probe_GPU2(const int* gpuCache, const probeArgsGPU* pargs, const offsetGPU* out_off,
    int num_tuples, int* total, int start_offset, const short* segment_group,
    mordred_stream_t stream) {
  INVALID_IF(!seg_ok(gpuCache, num_tuples, start_offset, segment_group) || !total || !out_off ||
             !out_off->lo_off);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.seg_group = segment_group; P.start_offset = start_offset; P.total = total;
  put_probe(P, pargs, 0xF);
  put_out_offsets(P, out_off);
  return dispatch(P, SEG, COMPACT, (hipStream_t)stream);
}

int mordred_probe_GPU3(const int* gpuCache, const offsetGPU* in_off, const probeArgsGPU* pargs,
    const offsetGPU* out_off, int num_tuples, int* total, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || !in_off || !in_off->lo_off || !total || !out_off || !out_off->lo_off);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.total = total;
  put_in_offsets(P, in_off);
  put_probe(P, pargs, 0xF);
  put_out_offsets(P, out_off);
  return dispatch(P, OFF, COMPACT, (hipStream_t)stream);
}

int mordred_filter_GPU2(const int* gpuCache, const filterArgsGPU* fargs, int* out_off,
    int num_tuples, int* total, int start_offset, const short* segment_group,
    mordred_stream_t stream) {
  INVALID_IF(!seg_ok(gpuCache, num_tuples, start_offset, segment_group) || !total || !out_off);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.seg_group = segment_group; P.start_offset = start_offset; P.total = total;
  put_filter_between(P, fargs, F_MODE);
  P.out[0] = out_off;
  return dispatch(P, SEG, COMPACT, (hipStream_t)stream);
}

int mordred_filter_GPU3(const int* gpuCache, const int* off_col, const filterArgsGPU* fargs,
    int* out_off, int num_tuples, int* total, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || !off_col || !total || !out_off);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.total = total;
  P.in_lo = off_col;
  P.aligned = aligned16(off_col);
  put_filter_between(P, fargs, F_MODE);
  P.out[0] = out_off;
  return dispatch(P, OFF, COMPACT, (hipStream_t)stream);
}

int mordred_aggregationGPU(const int* gpuCache, const int* lo_off, const groupbyArgsGPU* gargs,
    int num_tuples, int* res, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || !lo_off || !res);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.res = res;
  P.in_lo = lo_off;
  P.aligned = aligned16(lo_off);
  put_groupby(P, gargs);
  return dispatch(P, OFF, AGGR, (hipStream_t)stream);
}

int mordred_probe_aggr_GPU2(const int* gpuCache, const probeArgsGPU* pargs,
    const groupbyArgsGPU* gargs, int num_tuples, int* res, int start_offset,
    const short* segment_group, mordred_stream_t stream) {
  return mordred_filter_probe_aggr_GPU2(gpuCache, nullptr, pargs, gargs, num_tuples, res,
                                        start_offset, segment_group, stream);
}

int mordred_probe_aggr_GPU3(const int* gpuCache, const offsetGPU* offset, const probeArgsGPU* pargs,
    const groupbyArgsGPU* gargs, int num_tuples, int* res, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || !offset || !offset->lo_off || !res);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.res = res;
  put_in_offsets(P, offset);
  put_probe(P, pargs, 0x8);
  put_groupby(P, gargs);
  return dispatch(P, OFF, AGGR, (hipStream_t)stream);
}

int mordred_filter_probe_aggr_GPU2(const int* gpuCache, const filterArgsGPU* fargs,
    const probeArgsGPU* pargs, const groupbyArgsGPU* gargs, int num_tuples, int* res,
    int start_offset, const short* segment_group, mordred_stream_t stream) {
  INVALID_IF(!seg_ok(gpuCache, num_tuples, start_offset, segment_group) || !res);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.seg_group = segment_group; P.start_offset = start_offset; P.res = res;
  put_filter_between(P, fargs, F_BETWEEN);
  put_probe(P, pargs, 0x8);
  put_groupby(P, gargs);
  return dispatch(P, SEG, AGGR, (hipStream_t)stream);
}

int mordred_filter_probe_aggr_GPU3(const int* gpuCache, const offsetGPU* offset,
    const filterArgsGPU* fargs, const probeArgsGPU* pargs, const groupbyArgsGPU* gargs,
    int num_tuples, int* res, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || !offset || !offset->lo_off || !res);
  Pipe P = make_pipe(gpuCache, num_tuples);
  P.res = res;
  put_in_offsets(P, offset);
  put_filter_between(P, fargs, F_BETWEEN2ONLY);
  put_probe(P, pargs, 0x8);
  put_groupby(P, gargs);
  return dispatch(P, OFF, AGGR, (hipStream_t)stream);
}

// ------------------------------------------------------------------ builds

int mordred_build_GPU2(const int* gpuCache, const filterArgsGPU* fargs, const buildArgsGPU* bargs,
    int num_tuples, int* hash_table, int start_offset, const short* segment_group,
    mordred_stream_t stream) {
  INVALID_IF(!seg_ok(gpuCache, num_tuples, start_offset, segment_group) || !bargs ||
             !bargs->key_idx || !hash_table || bargs->num_slots <= 0);
  Build B{};
  B.cache = gpuCache; B.seg_group = segment_group; B.start_offset = start_offset;
  B.num_tuples = num_tuples; B.aligned = aligned16(gpuCache);
  if (fargs) { B.fcol = fargs->filter_idx1; B.fmode = fargs->mode1; B.c1 = fargs->compare1; B.c2 = fargs->compare2; }
  B.key = bargs->key_idx; B.val = bargs->val_idx; B.ht = hash_table;
  B.num_slots = bargs->num_slots; B.key_min = bargs->val_min;
  return run_build<SEG>(B, (hipStream_t)stream);
}

int mordred_build_GPU_minmax(const int* gpuCache, const filterArgsGPU* fargs, const buildArgsGPU* bargs,
    int num_tuples, int* hash_table, int* min_global, int* max_global, int start_offset,
    const short* segment_group, mordred_stream_t stream) {
  INVALID_IF(!seg_ok(gpuCache, num_tuples, start_offset, segment_group) || !bargs ||
             !bargs->key_idx || !hash_table || bargs->num_slots <= 0 || !min_global || !max_global);
  Build B{};
  B.cache = gpuCache; B.seg_group = segment_group; B.start_offset = start_offset;
  B.num_tuples = num_tuples; B.aligned = aligned16(gpuCache);
  if (fargs) { B.fcol = fargs->filter_idx1; B.fmode = fargs->mode1; B.c1 = fargs->compare1; B.c2 = fargs->compare2; }
  B.key = bargs->key_idx; B.val = bargs->val_idx; B.ht = hash_table;
  B.num_slots = bargs->num_slots; B.key_min = bargs->val_min;
  B.kmin_out = min_global; B.kmax_out = max_global;
  B.kmin_init = bargs->val_max; B.kmax_init = bargs->val_min;
  return run_build<SEG>(B, (hipStream_t)stream);
}

int mordred_build_GPU_minmax2(const int* gpuCache, const int* dim_off, const filterArgsGPU* fargs,
    const buildArgsGPU* bargs, int num_tuples, int* hash_table, int* min_global, int* max_global,
    mordred_stream_t stream) {
  // build_GPU_minmax2 asserts that no filter is given (GPUProcessing.h:1141).
  INVALID_IF(num_tuples < 0 || !dim_off || !bargs || !bargs->key_idx || !hash_table ||
             bargs->num_slots <= 0 || (fargs && fargs->filter_idx1) || !min_global || !max_global);
  Build B{};
  B.cache = gpuCache; B.num_tuples = num_tuples; B.in_rows = dim_off; B.aligned = aligned16(dim_off);
  B.key = bargs->key_idx; B.val = bargs->val_idx; B.ht = hash_table;
  B.num_slots = bargs->num_slots; B.key_min = bargs->val_min;
  B.kmin_out = min_global; B.kmax_out = max_global;
  B.kmin_init = bargs->val_max; B.kmax_init = bargs->val_min;
  return run_build<OFF>(B, (hipStream_t)stream);
}

int mordred_build_GPU3(const int* gpuCache, const int* dim_off, const filterArgsGPU* fargs,
    const buildArgsGPU* bargs, int num_tuples, int* hash_table, mordred_stream_t stream) {
  // build_GPU3 asserts that no filter is given (GPUProcessing.h:1017).
  INVALID_IF(num_tuples < 0 || !dim_off || !bargs || !bargs->key_idx || !hash_table ||
             bargs->num_slots <= 0 || (fargs && fargs->filter_idx1));
  Build B{};
  B.cache = gpuCache; B.num_tuples = num_tuples; B.in_rows = dim_off; B.aligned = aligned16(dim_off);
  B.key = bargs->key_idx; B.val = bargs->val_idx; B.ht = hash_table;
  B.num_slots = bargs->num_slots; B.key_min = bargs->val_min;
  return run_build<OFF>(B, (hipStream_t)stream);
}

int mordred_build_GPU(const int* key, const int* val, const int* filter, int compare1,
    int compare2, int filter_mode, int num_tuples, int* hash_table, int num_slots, int val_min,
    int segment_index, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || (num_tuples > 0 && !key) || !hash_table || num_slots <= 0);
  Build B{};
  B.num_tuples = num_tuples; B.row_base = segment_index * kSegSize;
  B.aligned = aligned16(key) && aligned16(val) && aligned16(filter);
  B.fcol = filter; B.fmode = filter_mode; B.c1 = compare1; B.c2 = compare2;
  B.key = key; B.val = val; B.ht = hash_table; B.num_slots = num_slots; B.key_min = val_min;
  return run_build<PLAIN>(B, (hipStream_t)stream);
}

int mordred_filter_probe_aggr_GPU(const int* filter1, const int* filter2, int compare1,
    int compare2, int compare3, int compare4, int filter_mode1, int filter_mode2, const int* key4,
    const int* ht4, int dim_len4, int min_key4, const int* aggr1, const int* aggr2,
    int num_tuples, int* res, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || !res);
  Pipe P = make_pipe(nullptr, num_tuples);
  P.res = res;
  P.fkind = F_MODE;
  P.fcol[0] = filter1; P.fcol[1] = filter2;
  P.cmp[0] = compare1; P.cmp[1] = compare2; P.cmp[2] = compare3; P.cmp[3] = compare4;
  P.fmode[0] = filter_mode1; P.fmode[1] = filter_mode2;
  P.pmask = 0x8;
  P.key[3] = key4; P.ht[3] = ht4; P.len[3] = dim_len4; P.minkey[3] = min_key4;
  P.aggr[0] = aggr1; P.aggr[1] = aggr2;
  P.aligned = aligned16(filter1) && aligned16(filter2) && aligned16(key4) && aligned16(aggr1) &&
              aligned16(aggr2);
  return dispatch(P, PLAIN, AGGR, (hipStream_t)stream);
}

int mordred_probe_group_by_GPU(const int* key1, const int* key2, const int* key3, const int* key4,
    const int* ht1, int dim_len1, const int* ht2, int dim_len2, const int* ht3, int dim_len3,
    const int* ht4, int dim_len4, int min_key1, int min_key2, int min_key3, int min_key4,
    const int* aggr1, const int* aggr2, int min_val1, int min_val2, int min_val3, int min_val4,
    int unique_val1, int unique_val2, int unique_val3, int unique_val4, int total_val,
    int num_tuples, int* res, mordred_stream_t stream) {
  INVALID_IF(num_tuples < 0 || !res || total_val <= 0);
  Pipe P = make_pipe(nullptr, num_tuples);
  P.res = res;
  P.pmask = 0xF;
  P.key[0] = key1; P.key[1] = key2; P.key[2] = key3; P.key[3] = key4;
  P.ht[0] = ht1; P.ht[1] = ht2; P.ht[2] = ht3; P.ht[3] = ht4;
  P.len[0] = dim_len1; P.len[1] = dim_len2; P.len[2] = dim_len3; P.len[3] = dim_len4;
  P.minkey[0] = min_key1; P.minkey[1] = min_key2; P.minkey[2] = min_key3; P.minkey[3] = min_key4;
  P.aggr[0] = aggr1; P.aggr[1] = aggr2;
  P.minv[0] = min_val1; P.minv[1] = min_val2; P.minv[2] = min_val3; P.minv[3] = min_val4;
  P.uniq[0] = unique_val1; P.uniq[1] = unique_val2; P.uniq[2] = unique_val3; P.uniq[3] = unique_val4;
  P.total_val = total_val;
  P.aligned = aligned16(key1) && aligned16(key2) && aligned16(key3) && aligned16(key4) &&
              aligned16(aggr1) && aligned16(aggr2);
  return dispatch(P, PLAIN, kGroupBySink, (hipStream_t)stream);
}

}  // extern "C"
