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

#include "../../include/mordred_hip.h"
#include "tile.hpp"

namespace mordred {

enum Mode { SEG = 0, OFF = 1, PLAIN = 2 };
enum Sink { GROUPBY_GLOBAL = 0, GROUPBY_LDS = 1, AGGR = 2, COMPACT = 3 };
enum FilterKind {
  F_NONE = 0,
  F_BETWEEN = 1,       // fused kernels: filter1 in [c1,c2], filter2 in [c3,c4]
  F_BETWEEN2ONLY = 2,  // filter_probe_aggr_GPU3 re-checks filter2 only (GPUProcessing.h:1805)
  F_MODE = 3           // filter kernels: mode 0 EQ, 1 BETWEEN, 2 EQ-or-EQ
};

struct Pipe {
  const int* cache;
  const short* seg_group;
  int start_offset;
  int num_tuples;
  int row_base;   // PLAIN: row id of position 0
  int aligned;    // OFF/PLAIN: all contiguous inputs are 16-byte aligned
  const int* in_lo;
  const int* in_dim[4];
  int fkind;
  const int* fcol[2];
  int cmp[4];
  int fmode[2];
  int pmask;      // bit k: table k+1 takes part in probing / dim-offset output
  const int* key[4];
  const int* ht[4];
  int len[4];
  int minkey[4];
  const int* aggr[2];
  const int* grp[4];
  int minv[4];
  int uniq[4];
  int total_val;
  int* res;
  int* total;
  int* out[5];
};

struct Build {
  const int* cache;
  const short* seg_group;
  int start_offset;
  int num_tuples;
  int row_base;
  int aligned;
  const int* in_rows;
  const int* fcol;
  int fmode;
  int c1, c2;
  const int* key;
  const int* val;
  int* ht;
  int num_slots;
  int key_min;
};

// Contiguous tile pointer of a column: SEG -> cached segment run, PLAIN -> column.
template <int MODE>
__device__ __forceinline__ const int* tile_ptr(const int* cache, const int* col, int seg,
                                               int start) {
  if (MODE == SEG) return cache + ((size_t)col[seg] << kSegLog) + (start & kSegMask);
  return col + start;
}

// Column values of this thread's items.  OFF mode gathers only items in `m`.
template <int BLOCK, int VPT, int MODE>
__device__ __forceinline__ void load_col(const int* cache, const int* col, int seg, int start,
                                         bool vec, int n, const int (&lo)[4 * VPT], uint32_t m,
                                         int (&x)[4 * VPT]) {
  if (MODE == OFF) {
#pragma unroll
    for (int j = 0; j < 4 * VPT; j++) x[j] = (m >> j & 1u) ? gather(cache, col, lo[j]) : 0;
  } else {
    load_tile<BLOCK, VPT>(tile_ptr<MODE>(cache, col, seg, start), vec, n, x);
  }
}

template <int VPT>
__device__ __forceinline__ uint32_t pred_between(const int (&x)[4 * VPT], int lo, int hi) {
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < 4 * VPT; j++) m |= (uint32_t)(x[j] >= lo && x[j] <= hi) << j;
  return m;
}

template <int VPT>
__device__ __forceinline__ uint32_t pred_eq(const int (&x)[4 * VPT], int c) {
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < 4 * VPT; j++) m |= (uint32_t)(x[j] == c) << j;
  return m;
}

// Mode predicate of filter_GPU2 / build_GPU2 (GPUProcessing.h:1231-1239, 972-980).
template <int VPT>
__device__ __forceinline__ uint32_t pred_mode(const int (&x)[4 * VPT], int mode, int c1, int c2,
                                              uint32_t all) {
  if (mode == 0) return pred_eq<VPT>(x, c1);
  if (mode == 1) return pred_between<VPT>(x, c1, c2);
  if (mode == 2) return pred_eq<VPT>(x, c1) | pred_eq<VPT>(x, c2);
  return all;
}

template <int BLOCK, int VPT, int MODE, int SINK>
__global__ __launch_bounds__(BLOCK) void pipeline_kernel(Pipe P) {
  constexpr int IPT = 4 * VPT;
  constexpr int TILE = BLOCK * IPT;
  __shared__ long long red[BLOCK / kWave];
  __shared__ int scan[BLOCK / kWave + 1];
  extern __shared__ unsigned long long gsum[];  // GROUPBY_LDS: sums, then touch bitmap
  uint32_t* touched = reinterpret_cast<uint32_t*>(gsum + (SINK == GROUPBY_LDS ? P.total_val : 0));

  if (SINK == GROUPBY_LDS) {
    const int words = P.total_val + (P.total_val + 63) / 64;  // sums + bitmap (as u64 words)
    for (int i = threadIdx.x; i < words; i += BLOCK) gsum[i] = 0;
    __syncthreads();
  }
  long long acc = 0;
  const int ntiles = (P.num_tuples + TILE - 1) / TILE;

  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int start = tile * TILE;
    const int n = min(TILE, P.num_tuples - start);
    const bool full = n == TILE;
    const uint32_t rmask = full ? (IPT == 32 ? 0xffffffffu : (1u << IPT) - 1u)
                                : range_mask<BLOCK, VPT>(n);
    const bool vec = full && P.aligned;
    uint32_t sel = rmask;

    int seg = 0, base_row = 0;
    int lo[IPT];
    if (MODE == SEG) {
      seg = P.seg_group ? (int)P.seg_group[start >> kSegLog] : (P.start_offset + start) >> kSegLog;
      base_row = seg * kSegSize + (start & kSegMask);
    } else if (MODE == PLAIN) {
      base_row = P.row_base + start;
    }
    if (MODE == OFF) {
      load_tile<BLOCK, VPT>(P.in_lo + start, vec, n, lo);
    } else {
#pragma unroll
      for (int j = 0; j < IPT; j++) lo[j] = base_row + item_row<BLOCK>(j);
    }

    int x[IPT];
    // ---------------------------------------------------------- filters
    if (P.fkind == F_BETWEEN || P.fkind == F_BETWEEN2ONLY) {
      if (P.fkind == F_BETWEEN && P.fcol[0]) {
        load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[0], seg, start, vec, n, lo, sel, x);
        sel &= pred_between<VPT>(x, P.cmp[0], P.cmp[1]);
      }
      if (P.fcol[1] && wave_any(sel)) {
        load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[1], seg, start, vec, n, lo, sel, x);
        sel &= pred_between<VPT>(x, P.cmp[2], P.cmp[3]);
      }
    } else if (P.fkind == F_MODE) {
      if (P.fcol[0]) {
        load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[0], seg, start, vec, n, lo, sel, x);
        sel &= pred_mode<VPT>(x, P.fmode[0], P.cmp[0], P.cmp[1], rmask);
      }
      if (P.fcol[1]) {
        if (P.fmode[1] == 2) {
          // crystal AndEQ then OrEQ: (sel && x==c3) || x==c4 (GPUProcessing.h:1252-1253)
          load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[1], seg, start, vec, n, lo, rmask, x);
          sel = ((sel & pred_eq<VPT>(x, P.cmp[2])) | pred_eq<VPT>(x, P.cmp[3])) & rmask;
        } else if (wave_any(sel)) {
          load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[1], seg, start, vec, n, lo, sel, x);
          if (P.fmode[1] == 0) sel &= pred_eq<VPT>(x, P.cmp[2]);
          else if (P.fmode[1] == 1) sel &= pred_between<VPT>(x, P.cmp[2], P.cmp[3]);
        }
      }
    }

    // ---------------------------------------------------------- probes
    int dv[4][IPT];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int dflt = SINK == COMPACT ? 1 : 0;  // BlockSetValue(...,1) / (...,0)
#pragma unroll
      for (int j = 0; j < IPT; j++) dv[k][j] = dflt;
      if (!(P.pmask >> k & 1)) continue;
      if (P.key[k] && P.ht[k]) {
        if (!wave_any(sel)) continue;
        load_col<BLOCK, VPT, MODE>(P.cache, P.key[k], seg, start, vec, n, lo, sel, x);
        const int* ht = P.ht[k];
        const int len = P.len[k], mk = P.minkey[k];
#pragma unroll
        for (int j = 0; j < IPT; j++) {
          if (!(sel >> j & 1u)) continue;
          const int h = phash(x[j], len, mk);
          if (SINK == COMPACT) {
            // BlockProbeGPU: hit iff ht[2h+1] != 0, dimension row = slot-1
            const int s = ht[2 * (size_t)h + 1];
            if (s != 0) dv[k][j] = s - 1;
            else sel &= ~(1u << j);
          } else {
            // BlockProbeGroupByGPU: hit iff the 64-bit slot != 0, value = low word
            const unsigned long long s =
                *reinterpret_cast<const unsigned long long*>(ht + 2 * (size_t)h);
            if (s != 0ull) dv[k][j] = (int)(unsigned)s;
            else sel &= ~(1u << j);
          }
        }
      } else if (MODE == OFF && SINK == COMPACT && P.in_dim[k]) {
        // pass an earlier probe's dimension row through (GPUProcessing.h:859-861)
        load_tile<BLOCK, VPT>(P.in_dim[k] + start, vec, n, dv[k]);
      } else if (MODE == OFF && (SINK == GROUPBY_LDS || SINK == GROUPBY_GLOBAL) && P.grp[k]) {
        // group value from an earlier probe's dimension row (GPUProcessing.h:203-206)
        int d[IPT];
        load_tile<BLOCK, VPT>(P.in_dim[k] + start, vec, n, d);
#pragma unroll
        for (int j = 0; j < IPT; j++)
          if (sel >> j & 1u) dv[k][j] = gather(P.cache, P.grp[k], d[j]);
      }
    }

    // ---------------------------------------------------------- sinks
    if (SINK == COMPACT) {
      int pos = block_reserve<BLOCK>(__popc(sel), P.total, scan);
#pragma unroll
      for (int j = 0; j < IPT; j++) {
        if (!(sel >> j & 1u)) continue;
        P.out[0][pos] = lo[j];
#pragma unroll
        for (int k = 0; k < 4; k++)
          if ((P.pmask >> k & 1) && P.out[k + 1]) P.out[k + 1][pos] = dv[k][j];
        pos++;
      }
      continue;
    }
    if (!wave_any(sel)) continue;
    int a1[IPT], a2[IPT];
    if (P.aggr[0]) load_col<BLOCK, VPT, MODE>(P.cache, P.aggr[0], seg, start, vec, n, lo, sel, a1);
    else for (int j = 0; j < IPT; j++) a1[j] = 0;
    if (P.aggr[1]) load_col<BLOCK, VPT, MODE>(P.cache, P.aggr[1], seg, start, vec, n, lo, sel, a2);
    else for (int j = 0; j < IPT; j++) a2[j] = 0;

    if (SINK == AGGR) {
#pragma unroll
      for (int j = 0; j < IPT; j++)
        if (sel >> j & 1u) acc += (long long)(a1[j] * a2[j]);  // int32 product (GPUProcessing.h:1763)
      continue;
    }
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      if (!(sel >> j & 1u)) continue;
      const int h = ((dv[0][j] - P.minv[0]) * P.uniq[0] + (dv[1][j] - P.minv[1]) * P.uniq[1] +
                     (dv[2][j] - P.minv[2]) * P.uniq[2] + (dv[3][j] - P.minv[3]) * P.uniq[3]) %
                    P.total_val;
      const unsigned long long v = (unsigned long long)(long long)(a1[j] - a2[j]);
      int2* key = reinterpret_cast<int2*>(P.res + 6 * (size_t)h);
      if (SINK == GROUPBY_GLOBAL) {
        key[0] = make_int2(dv[0][j], dv[1][j]);
        key[1] = make_int2(dv[2][j], dv[3][j]);
        atomicAdd(reinterpret_cast<unsigned long long*>(key + 2), v);
      } else {
        const uint32_t bit = 1u << (h & 31);
        if (!(atomicOr(&touched[h >> 5], bit) & bit)) {
          key[0] = make_int2(dv[0][j], dv[1][j]);
          key[1] = make_int2(dv[2][j], dv[3][j]);
        }
        atomicAdd(&gsum[h], v);
      }
    }
  }

  if (SINK == GROUPBY_LDS) {
    __syncthreads();
    for (int h = threadIdx.x; h < P.total_val; h += BLOCK)
      if (touched[h >> 5] >> (h & 31) & 1u)
        atomicAdd(reinterpret_cast<unsigned long long*>(P.res + 6 * (size_t)h + 4), gsum[h]);
  } else if (SINK == AGGR) {
    const long long s = block_sum<BLOCK>(acc, red);
    if (threadIdx.x == 0 && s != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(P.res + 4), (unsigned long long)s);
  }
}

// Perfect-hash build: BlockBuildValueGPU / BlockBuildOffsetGPU
// (BlockLibrary.cuh:594-735) and their offset-list twins (:737-890).
template <int BLOCK, int VPT, int MODE>
__global__ __launch_bounds__(BLOCK) void build_kernel(Build B) {
  constexpr int IPT = 4 * VPT;
  constexpr int TILE = BLOCK * IPT;
  const int ntiles = (B.num_tuples + TILE - 1) / TILE;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int start = tile * TILE;
    const int n = min(TILE, B.num_tuples - start);
    const bool full = n == TILE;
    const uint32_t rmask = full ? (IPT == 32 ? 0xffffffffu : (1u << IPT) - 1u)
                                : range_mask<BLOCK, VPT>(n);
    const bool vec = full && B.aligned;
    uint32_t sel = rmask;
    int seg = 0, base_row = 0, row[IPT];
    if (MODE == SEG) {
      seg = B.seg_group ? (int)B.seg_group[start >> kSegLog] : (B.start_offset + start) >> kSegLog;
      base_row = seg * kSegSize + (start & kSegMask);
    } else if (MODE == PLAIN) {
      base_row = B.row_base + start;
    }
    if (MODE == OFF) {
      load_tile<BLOCK, VPT>(B.in_rows + start, vec, n, row);
    } else {
#pragma unroll
      for (int j = 0; j < IPT; j++) row[j] = base_row + item_row<BLOCK>(j);
    }
    int x[IPT], v[IPT];
    if (MODE != OFF && B.fcol) {
      load_col<BLOCK, VPT, MODE>(B.cache, B.fcol, seg, start, vec, n, row, sel, x);
      sel &= pred_mode<VPT>(x, B.fmode, B.c1, B.c2, rmask);
    }
    if (!wave_any(sel)) continue;
    load_col<BLOCK, VPT, MODE>(B.cache, B.key, seg, start, vec, n, row, sel, x);
    if (B.val) load_col<BLOCK, VPT, MODE>(B.cache, B.val, seg, start, vec, n, row, sel, v);
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      if (!(sel >> j & 1u)) continue;
      const size_t h = (size_t)phash(x[j], B.num_slots, B.key_min);
      if (B.val) atomicCAS(&B.ht[2 * h], 0, v[j]);
      B.ht[2 * h + 1] = row[j] + 1;
    }
  }
}

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
static int resident_groups(const void* f, int block, size_t lds) {
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

constexpr int kBlock = 256;
constexpr int kVpt = 2;
constexpr int kLdsBlock = 512;
constexpr size_t kLdsGroupBudget = 96 * 1024;

static size_t lds_bytes_for(int total_val) {
  return ((size_t)total_val + ((size_t)total_val + 63) / 64) * sizeof(unsigned long long);
}

template <int BLOCK, int VPT, int MODE, int SINK>
static int launch_pipe(const Pipe& P, hipStream_t stream, size_t lds) {
  constexpr int TILE = BLOCK * 4 * VPT;
  const void* f = reinterpret_cast<const void*>(&pipeline_kernel<BLOCK, VPT, MODE, SINK>);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const int ntiles = (P.num_tuples + TILE - 1) / TILE;
  const int grid = std::max(1, std::min(ntiles, resident_groups(f, BLOCK, lds)));
  pipeline_kernel<BLOCK, VPT, MODE, SINK><<<grid, BLOCK, lds, stream>>>(P);
  return hipGetLastError();
}

template <int MODE>
static int run_pipe(const Pipe& P, int sink, hipStream_t stream) {
  switch (sink) {
    case GROUPBY_LDS:
      return launch_pipe<kLdsBlock, kVpt, MODE, GROUPBY_LDS>(P, stream, lds_bytes_for(P.total_val));
    case GROUPBY_GLOBAL: return launch_pipe<kBlock, kVpt, MODE, GROUPBY_GLOBAL>(P, stream, 0);
    case AGGR: return launch_pipe<kBlock, kVpt, MODE, AGGR>(P, stream, 0);
    default: return launch_pipe<kBlock, kVpt, MODE, COMPACT>(P, stream, 0);
  }
}

static int dispatch(const Pipe& P, int mode, int sink, hipStream_t stream) {
  if (P.num_tuples == 0) return hipSuccess;
  if (mode == SEG) return run_pipe<SEG>(P, sink, stream);
  if (mode == OFF) return run_pipe<OFF>(P, sink, stream);
  return run_pipe<PLAIN>(P, sink, stream);
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

static int groupby_sink(const Pipe& P) {
  return lds_bytes_for(P.total_val) <= kLdsGroupBudget ? GROUPBY_LDS : GROUPBY_GLOBAL;
}

template <int MODE>
static int run_build(const Build& B, hipStream_t stream) {
  if (B.num_tuples == 0) return hipSuccess;
  constexpr int TILE = kBlock * 4 * kVpt;
  const void* f = reinterpret_cast<const void*>(&build_kernel<kBlock, kVpt, MODE>);
  const int ntiles = (B.num_tuples + TILE - 1) / TILE;
  const int grid = std::max(1, std::min(ntiles, resident_groups(f, kBlock, 0)));
  build_kernel<kBlock, kVpt, MODE><<<grid, kBlock, 0, stream>>>(B);
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
  return dispatch(P, SEG, groupby_sink(P), (hipStream_t)stream);
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
  return dispatch(P, OFF, groupby_sink(P), (hipStream_t)stream);
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
  return dispatch(P, PLAIN, groupby_sink(P), (hipStream_t)stream);
}

}  // extern "C"
