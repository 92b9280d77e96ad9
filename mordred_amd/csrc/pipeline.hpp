// Device side of the query pipeline: argument blocks, the generic tile
// pipeline kernel and the perfect-hash build kernel.  Included by every
// translation unit that instantiates kernels (kernels.hip for the generic
// shapes, spec_*.hip for the compile-time specialised query shapes).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/mordred_hip.h"
#include "probe_tables.hpp"
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
  // kdecode: the group index is a mixed-radix number with digit k in
  // [0, krange[k]) for uniq[k] != 0, so keys in range decode from the index
  int krange[4];
  int kdecode;
  int total_val;
  int* res;
  int* total;
  int* out[5];
  // compact probe tables (probe_tables.hpp): header + packed words per table,
  // nullptr = probe the hash table itself
  const CompactHdr* chdr[4];
  const uint32_t* cwords[4];
  const uint32_t* cbits[4];  // existence bitmaps of large compact tables, or nullptr
  // Compact tables in `lds_stage` (bit k = table k) are copied to dynamic LDS
  // at the start of the kernel, in probe order, while their packed words fit
  // in `lds_words` 32-bit words (entry widths are known only on the device).
  int lds_stage;
  int lds_words;
  int ring_off;      // u64 offset of the survivor rings in dynamic LDS
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
  // build_GPU_minmax / build_GPU_minmax2: min / max of the built keys merged
  // into *kmin_out / *kmax_out with atomics; every workgroup starts from
  // (kmin_init, kmax_init) = (val_max, val_min) as GPUProcessing.h:1050 does.
  // nullptr: a plain build.
  int* kmin_out;
  int* kmax_out;
  int kmin_init, kmax_init;
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

// Segment mode: the cache slot of each column of the current segment, loaded
// once per segment (all columns at once, made wave-uniform) instead of once
// per column per tile: a slot load followed by a wait would also wait for the
// data loads of the columns issued before it (loads complete in issue order).
template <int N>
struct SegSlots {
  int gi = -1;       // segment-group position of the cached slots
  int seg = 0;
  int slot[N];
};

template <int N>
__device__ __forceinline__ void seg_slots(SegSlots<N>& S, const short* seg_group, int start_offset, int start,
                                          const int* const (&cols)[N], unsigned used) {
  const int gi = start >> kSegLog;
  if (gi == S.gi) return;
  S.gi = gi;
  const int seg = seg_group ? (int)seg_group[gi] : (start_offset + start) >> kSegLog;
  S.seg = __builtin_amdgcn_readfirstlane(seg);
  int v[N];
#pragma unroll
  for (int c = 0; c < N; c++) v[c] = (used >> c & 1u) ? cols[c][S.seg] : 0;
#pragma unroll
  for (int c = 0; c < N; c++) S.slot[c] = __builtin_amdgcn_readfirstlane(v[c]);
}

// Contiguous tile range of this workgroup (persistent grid of balanced_grid
// size): consecutive tiles share a segment.
__device__ __forceinline__ void tile_range(int ntiles, int& t0, int& t1) {
  const int per = (ntiles + (int)gridDim.x - 1) / (int)gridDim.x;
  t0 = (int)blockIdx.x * per;
  t1 = min(ntiles, t0 + per);
}

// Tile load of a contiguous (SEG / PLAIN) tile; SEG takes the column's slot.
template <int BLOCK, int VPT, int MODE>
__device__ __forceinline__ void load_eager(const int* cache, const int* col, int slot, int start, bool vec,
                                           int n, int (&x)[4 * VPT]) {
  static_assert(MODE != OFF, "contiguous tiles only");
  if (MODE == SEG) load_tile<BLOCK, VPT>(cache + ((size_t)slot << kSegLog) + (start & kSegMask), vec, n, x);
  else load_tile<BLOCK, VPT>(col + start, vec, n, x);
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

// SHAPE < 0: everything decided at run time from the argument pointers.
// SHAPE >= 0 (segment mode, fact BETWEEN filters, group-by/aggregate sinks):
// bits 0-3 = tables probed, bit 4 / bit 5 = filter column 1 / 2 present; the
// code for absent tables and filters is compiled out.
constexpr int kShapeGeneric = -1;

// Survivor compaction (specialised segment-mode group-by / aggregate shapes
// that probe two or more tables): after the first probe and the existence-
// bitmap tests of later tables, each wave appends the surviving rows to a
// ring in LDS (positions from ballot + mbcnt): the row id, the first probe's
// result and the keys already loaded.  The later probes and the sink run on
// full rounds of 64 dense entries instead of on every item slot, and gather
// the remaining keys and the aggregates by row (late materialization: only
// cache lines that hold a survivor are read).
constexpr int kRing = 128;   // entries per wave (a round of 64 plus one item of carry)
// Ring fields of a shape: one per probed table (its value or key), then the
// row id; table k of the shape uses field ring_field(shape, k).
constexpr int popc4(int s) { return (s & 1) + (s >> 1 & 1) + (s >> 2 & 1) + (s >> 3 & 1); }
constexpr int ring_fields(int shape) { return popc4(shape & 0xF) + 1; }
constexpr int ring_field(int shape, int k) { return popc4(shape & 0xF & ((1 << k) - 1)); }
constexpr bool shape_compacts(int mode, int sink, int shape) {
  return shape >= 0 && mode == SEG && (sink == GROUPBY_LDS || sink == AGGR) &&
         ((shape & 1) + (shape >> 1 & 1) + (shape >> 2 & 1) + (shape >> 3 & 1)) >= 2;
}
constexpr int shape_first(int shape) {
  return (shape & 1) ? 0 : (shape & 2) ? 1 : (shape & 4) ? 2 : (shape & 8) ? 3 : -1;
}

// Whether keys v decode back from their group index: with P.kdecode, every
// grouping dimension's digit v[k] - minv[k] lies in [0, krange[k]) and every
// other key is 0.  Such keys are not stored per row; the flush writes them
// from the index.
__device__ __forceinline__ bool keys_decode(const Pipe& P, const int v[4]) {
  bool d = P.kdecode != 0;
#pragma unroll
  for (int k = 0; k < 4; k++)
    d = d && (P.uniq[k] != 0 ? (unsigned)v[k] - (unsigned)P.minv[k] < (unsigned)P.krange[k] : v[k] == 0);
  return d;
}

template <int BLOCK, int VPT, int MODE, int SINK, int SHAPE = kShapeGeneric>
__global__ __launch_bounds__(BLOCK) void pipeline_kernel(Pipe P) {
  constexpr bool kGen = SHAPE < 0;
  static_assert(kGen || (MODE != OFF && SINK != COMPACT), "specialised shapes: SEG/PLAIN group-by or aggregate");
  constexpr int IPT = 4 * VPT;
  constexpr int TILE = BLOCK * IPT;
  constexpr bool kGroupBy = SINK == GROUPBY_LDS || SINK == GROUPBY_GLOBAL;
  // SEG/PLAIN tiles are contiguous runs: every column load of a tile is issued
  // before any of them is used.  OFF tiles gather only rows still selected.
  constexpr bool kEager = MODE != OFF;
  __shared__ long long red[BLOCK / kWave];
  __shared__ int scan[BLOCK / kWave + 1];
  // dynamic LDS: [group sums (u64) | group touch bitmap | bitmap of groups
  // whose keys this workgroup stored] for GROUPBY_LDS, then the packed words
  // of the LDS-staged compact probe table
  extern __shared__ unsigned long long lds[];
  const int bwords = (P.total_val + 63) / 64;
  const int gwords = SINK == GROUPBY_LDS ? P.total_val + 2 * bwords : 0;
  unsigned long long* gsum = lds;
  uint32_t* touched = reinterpret_cast<uint32_t*>(lds + (SINK == GROUPBY_LDS ? P.total_val : 0));
  uint32_t* stored = touched + 2 * bwords;
  uint32_t* ldsw = reinterpret_cast<uint32_t*>(lds + gwords);
  constexpr bool kCompact = shape_compacts(MODE, SINK, SHAPE);
  constexpr int kFirst = shape_first(SHAPE);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  constexpr int kRowField = kCompact ? ring_fields(SHAPE) - 1 : 0;
  int* ring = kCompact ? reinterpret_cast<int*>(lds + P.ring_off) + wave * (kRing * ring_fields(SHAPE)) : nullptr;
  int ring_head = 0, ring_tail = 0;  // wave-uniform

  // per-table compact probe parameters (workgroup-uniform); cb: test table
  // k's existence bits per tile, before the ring, when fewer than half its
  // slots hit: the existence bitmap of a large table, or the packed words
  // themselves of a 1-bit table (a table without values, e.g. a filtered
  // part table that is not grouped on), which are that bitmap
  int cw[4], csh[4], cvmin[4];
  bool cb[4];
  const uint32_t* cbp[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    cw[k] = 0; csh[k] = 0; cvmin[k] = 0; cb[k] = false; cbp[k] = nullptr;
    if ((kGen || (SHAPE >> k & 1)) && P.chdr[k]) {
      const CompactHdr h = *P.chdr[k];
      cw[k] = h.w; csh[k] = h.shift; cvmin[k] = h.vmin;
      cbp[k] = h.w == 1 ? P.cwords[k] : P.cbits[k];
      cb[k] = cbp[k] != nullptr && h.w >= 1 && h.any < P.len[k] / 2;
    }
  }

  if (SINK == GROUPBY_LDS) {
    for (int i = threadIdx.x; i < gwords; i += BLOCK) lds[i] = 0;
  }
  // stage compact tables in LDS (workgroup-uniform decisions): table k's
  // packed words at ldsw + soff[k] when stg[k]; reads branch on stg[k] so
  // that they stay LDS or global loads (not flat)
  bool stg[4];
  int soff[4];
  bool any_staged = false;
  int cursor = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    stg[k] = false;
    soff[k] = 0;
    if (!((kGen || (SHAPE >> k & 1)) && (P.lds_stage >> k & 1) && cw[k] > 0)) continue;
    const int nwords = (P.len[k] + (1 << csh[k]) - 1) >> csh[k];
    if (cursor + nwords > P.lds_words) continue;
    stg[k] = true;
    soff[k] = cursor;
    any_staged = true;
    const uint32_t* src = P.cwords[k];
    for (int i = threadIdx.x; i < nwords; i += BLOCK) ldsw[cursor + i] = src[i];
    cursor += nwords;
  }
  if (SINK == GROUPBY_LDS || any_staged) __syncthreads();

  long long acc = 0;
  const int ntiles = (P.num_tuples + TILE - 1) / TILE;

  // one round over `count` ring entries from ring_head, one entry per lane:
  // the remaining probes, then the sink
  auto ring_round = [&](int count) {
    if constexpr (kCompact) {
      __builtin_amdgcn_wave_barrier();
      const int e = (ring_head + lane) & (kRing - 1);
      bool hit = lane < count;
      int v[4] = {0, 0, 0, 0};
#pragma unroll
      for (int k = 0; k < 4; k++)
        if ((SHAPE >> k & 1) && (k == kFirst || cb[k])) v[k] = ring[ring_field(SHAPE, k) * kRing + e];
      const int row = ring[kRowField * kRing + e];
      __builtin_amdgcn_wave_barrier();
      // late columns of the entry's row: segment slot, then value
      const int rseg = hit ? row >> kSegLog : 0, roff = row & kSegMask;
      int a0 = 0, a1 = 0;
#pragma unroll
      for (int k = 0; k < 4; k++)
        if ((SHAPE >> k & 1) && k != kFirst && !cb[k] && hit)
          v[k] = P.cache[((size_t)P.key[k][rseg] << kSegLog) + roff];
      if (P.aggr[0] && hit) a0 = P.cache[((size_t)P.aggr[0][rseg] << kSegLog) + roff];
      if (P.aggr[1] && hit) a1 = P.cache[((size_t)P.aggr[1][rseg] << kSegLog) + roff];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (!(SHAPE >> k & 1) || k == kFirst) continue;
        const int hs = hit ? phash(v[k], P.len[k], P.minkey[k]) : 0;
        if (cb[k] && cw[k] == 1) {
          // a 1-bit table: its bits were tested before the ring, and its
          // result is the constant vmin
          v[k] = cvmin[k];
        } else if (cw[k] > 0) {
          // (an existence bitmap, if any, was tested before the ring)
          const int w = cw[k], sh = csh[k];
          uint32_t word = 0;
          if (stg[k]) {
            if (hit) word = ldsw[soff[k] + (hs >> sh)];
          } else if (hit) {
            word = P.cwords[k][hs >> sh];
          }
          const uint32_t ent = (word >> ((hs & ((1 << sh) - 1)) * w)) & ((1u << w) - 1u);
          hit = hit && ent != 0u;
          v[k] = cvmin[k] + (int)ent - 1;
        } else {
          const unsigned long long sv = *reinterpret_cast<const unsigned long long*>(P.ht[k] + 2 * (size_t)hs);
          hit = hit && sv != 0ull;
          v[k] = (int)(unsigned)sv;
        }
      }
      if (SINK == AGGR) {
        if (hit) acc += (long long)(a0 * a1);  // int32 product (GPUProcessing.h:1763)
      } else if (hit) {
        const int g = mod_pos((v[0] - P.minv[0]) * P.uniq[0] + (v[1] - P.minv[1]) * P.uniq[1] +
                                  (v[2] - P.minv[2]) * P.uniq[2] + (v[3] - P.minv[3]) * P.uniq[3],
                              P.total_val);
        atomicAdd(&gsum[g], (unsigned long long)(long long)(a0 - a1));
        const uint32_t bit = 1u << (g & 31);
        if (!(atomicOr(&touched[g >> 5], bit) & bit) && !keys_decode(P, v)) {
          int2* key = reinterpret_cast<int2*>(P.res + 6 * (size_t)g);
          key[0] = make_int2(v[0], v[1]);
          key[1] = make_int2(v[2], v[3]);
          atomicOr(&stored[g >> 5], bit);
        }
      }
    }
  };

  // slots 0-1 filter columns, 2-5 probe keys, 6-7 aggregates: those the
  // tile loads below read
  SegSlots<8> S;
  const int* const scols[8] = {P.fcol[0], P.fcol[1], P.key[0], P.key[1], P.key[2], P.key[3], P.aggr[0], P.aggr[1]};
  unsigned sused = 0;
#pragma unroll
  for (int c = 0; c < 2; c++) {
    if (kGen ? P.fcol[c] != nullptr : (SHAPE >> (4 + c) & 1)) sused |= 1u << c;
    if (P.aggr[c] && SINK != COMPACT && !kCompact) sused |= 1u << (6 + c);
  }
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (kGen ? ((P.pmask >> k & 1) && P.key[k] && P.ht[k]) : (SHAPE >> k & 1)) sused |= 1u << (2 + k);
  int tile0, tile1;
  tile_range(ntiles, tile0, tile1);
  for (int tile = tile0; tile < tile1; tile++) {
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
      seg_slots(S, P.seg_group, P.start_offset, start, scols, sused);
      seg = S.seg;
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

    // column values: fv filters, kv probe keys (replaced in place by the probe
    // result: group value or dimension row), av aggregates
    int fv[2][IPT], kv[4][IPT], av[2][IPT];
    if constexpr (kEager) {
#pragma unroll
      for (int c = 0; c < 2; c++)
        if (kGen ? P.fcol[c] != nullptr : (SHAPE >> (4 + c) & 1))
          load_eager<BLOCK, VPT, MODE>(P.cache, P.fcol[c], S.slot[c], start, vec, n, fv[c]);
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (kGen ? ((P.pmask >> k & 1) && P.key[k] && P.ht[k]) : (SHAPE >> k & 1))
          if (!kCompact || k == kFirst || cb[k])
            load_eager<BLOCK, VPT, MODE>(P.cache, P.key[k], S.slot[2 + k], start, vec, n, kv[k]);
      // with survivor compaction, the other keys and the aggregates are
      // gathered by row in the ring rounds (late materialization)
      if (SINK != COMPACT && !kCompact) {
#pragma unroll
        for (int c = 0; c < 2; c++)
          if (P.aggr[c]) load_eager<BLOCK, VPT, MODE>(P.cache, P.aggr[c], S.slot[6 + c], start, vec, n, av[c]);
      }
    }

    // ---------------------------------------------------------- filters
    if (!kGen) {
      if (SHAPE >> 4 & 1) sel &= pred_between<VPT>(fv[0], P.cmp[0], P.cmp[1]);
      if (SHAPE >> 5 & 1) sel &= pred_between<VPT>(fv[1], P.cmp[2], P.cmp[3]);
    } else if (P.fkind == F_BETWEEN || P.fkind == F_BETWEEN2ONLY) {
      if (P.fkind == F_BETWEEN && P.fcol[0]) {
        if (!kEager) load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[0], seg, start, vec, n, lo, sel, fv[0]);
        sel &= pred_between<VPT>(fv[0], P.cmp[0], P.cmp[1]);
      }
      if (P.fcol[1] && (kEager || wave_any(sel))) {
        if (!kEager) load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[1], seg, start, vec, n, lo, sel, fv[1]);
        sel &= pred_between<VPT>(fv[1], P.cmp[2], P.cmp[3]);
      }
    } else if (P.fkind == F_MODE) {
      if (P.fcol[0]) {
        if (!kEager) load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[0], seg, start, vec, n, lo, sel, fv[0]);
        sel &= pred_mode<VPT>(fv[0], P.fmode[0], P.cmp[0], P.cmp[1], rmask);
      }
      if (P.fcol[1]) {
        if (P.fmode[1] == 2) {
          // crystal AndEQ then OrEQ: (sel && x==c3) || x==c4 (GPUProcessing.h:1252-1253)
          if (!kEager) load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[1], seg, start, vec, n, lo, rmask, fv[1]);
          sel = ((sel & pred_eq<VPT>(fv[1], P.cmp[2])) | pred_eq<VPT>(fv[1], P.cmp[3])) & rmask;
        } else if (kEager || wave_any(sel)) {
          if (!kEager) load_col<BLOCK, VPT, MODE>(P.cache, P.fcol[1], seg, start, vec, n, lo, sel, fv[1]);
          if (P.fmode[1] == 0) sel &= pred_eq<VPT>(fv[1], P.cmp[2]);
          else if (P.fmode[1] == 1) sel &= pred_between<VPT>(fv[1], P.cmp[2], P.cmp[3]);
        }
      }
    }

    // ---------------------------------------------------------- probes
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (kCompact && k != kFirst) {
        // a later table with an existence bitmap: test the bitmap here, for
        // all items of the tile at once, so that only rows whose slot is
        // filled reach the ring (whose rounds then skip the bitmap)
        if (!(SHAPE >> k & 1) || !cb[k]) continue;
        uint32_t bw[IPT];
        int hb[IPT];
#pragma unroll
        for (int j = 0; j < IPT; j++) hb[j] = (sel >> j & 1u) ? phash(kv[k][j], P.len[k], P.minkey[k]) : 0;
        if (stg[k] && cw[k] == 1) {
          // a 1-bit table staged in LDS
#pragma unroll
          for (int j = 0; j < IPT; j++) bw[j] = ldsw[soff[k] + (hb[j] >> 5)];
        } else {
#pragma unroll
          for (int j = 0; j < IPT; j++) {
            // masked: lanes without a survivor issue no address
            if (sel >> j & 1u) bw[j] = cbp[k][hb[j] >> 5];
            else bw[j] = 0xffffffffu;
          }
        }
#pragma unroll
        for (int j = 0; j < IPT; j++)
          if (!(bw[j] >> (hb[j] & 31) & 1u)) sel &= ~(1u << j);
        continue;
      }
      const bool probed = kGen ? ((P.pmask >> k & 1) && P.key[k] && P.ht[k]) : (SHAPE >> k & 1);
      if (!probed) {
        const int dflt = SINK == COMPACT ? 1 : 0;  // BlockSetValue(...,1) / (...,0)
#pragma unroll
        for (int j = 0; j < IPT; j++) kv[k][j] = dflt;
        if (!kGen || !(P.pmask >> k & 1)) continue;
        if (MODE == OFF && SINK == COMPACT && P.in_dim[k]) {
          // pass an earlier probe's dimension row through (GPUProcessing.h:859-861)
          load_tile<BLOCK, VPT>(P.in_dim[k] + start, vec, n, kv[k]);
        } else if (MODE == OFF && kGroupBy && P.grp[k]) {
          // group value from an earlier probe's dimension row (GPUProcessing.h:203-206)
          int d[IPT];
          load_tile<BLOCK, VPT>(P.in_dim[k] + start, vec, n, d);
#pragma unroll
          for (int j = 0; j < IPT; j++)
            if (sel >> j & 1u) kv[k][j] = gather(P.cache, P.grp[k], d[j]);
        }
        continue;
      }
      if (!kEager) {
        if (!wave_any(sel)) continue;
        load_col<BLOCK, VPT, MODE>(P.cache, P.key[k], seg, start, vec, n, lo, sel, kv[k]);
      }
      const int* ht = P.ht[k];
      const int len = P.len[k], mk = P.minkey[k];
      // All slot reads of the tile are issued before any is used, so the
      // 4*VPT probes of a thread overlap instead of paying one L2/MALL round
      // trip each.  Compact-table reads in global memory are masked to the
      // items still selected (lanes without one issue no address, which
      // unloads the texture address unit); hash-table and LDS reads of
      // unselected items read slot 0 (measured faster for them).
      int hs[IPT];
      uint32_t act = sel;
#pragma unroll
      for (int j = 0; j < IPT; j++) {
        const int h = phash(kv[k][j], len, mk);
        const bool a = sel >> j & 1u;
        hs[j] = a ? h : 0;
      }
      if (cw[k] > 0) {
        // compact probe table: entry e = 0 on a miss, else result - vmin + 1
        const int w = cw[k], sh = csh[k], vmin = cvmin[k];
        const uint32_t emask = (1u << w) - 1u;
        uint32_t wv[IPT];
        if (stg[k]) {
          const uint32_t* lw = ldsw + soff[k];
#pragma unroll
          for (int j = 0; j < IPT; j++) wv[j] = lw[hs[j] >> sh];
        } else {
          const uint32_t* cwk = P.cwords[k];
#pragma unroll
          for (int j = 0; j < IPT; j++) {
            if (act >> j & 1u) wv[j] = cwk[hs[j] >> sh];
            else wv[j] = 0u;
          }
        }
#pragma unroll
        for (int j = 0; j < IPT; j++) {
          const uint32_t e = (wv[j] >> ((hs[j] & ((1 << sh) - 1)) * w)) & emask;
          const bool hit = (act >> j & 1u) && e != 0u;
          kv[k][j] = hit ? vmin + (int)e - 1 : kv[k][j];
          if (!hit) act &= ~(1u << j);
        }
      } else if (SINK == COMPACT) {
        // BlockProbeGPU: hit iff ht[2h+1] != 0, dimension row = slot-1
        int sv[IPT];
#pragma unroll
        for (int j = 0; j < IPT; j++) sv[j] = ht[2 * (size_t)hs[j] + 1];
#pragma unroll
        for (int j = 0; j < IPT; j++) {
          const bool hit = (act >> j & 1u) && sv[j] != 0;
          kv[k][j] = hit ? sv[j] - 1 : kv[k][j];
          if (!hit) act &= ~(1u << j);
        }
      } else {
        // BlockProbeGroupByGPU: hit iff the 64-bit slot != 0, value = low word
        unsigned long long sv[IPT];
#pragma unroll
        for (int j = 0; j < IPT; j++)
          sv[j] = *reinterpret_cast<const unsigned long long*>(ht + 2 * (size_t)hs[j]);
#pragma unroll
        for (int j = 0; j < IPT; j++) {
          const bool hit = (act >> j & 1u) && sv[j] != 0ull;
          kv[k][j] = hit ? (int)(unsigned)sv[j] : kv[k][j];
          if (!hit) act &= ~(1u << j);
        }
      }
      sel = act;
    }

    if constexpr (kCompact) {
      // append this tile's survivors to the wave's ring; run full rounds
#pragma unroll
      for (int j = 0; j < IPT; j++) {
        const unsigned long long b = __ballot(sel >> j & 1u);
        if (b == 0ull) continue;
        if (sel >> j & 1u) {
          const int pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
          const int e = (ring_tail + pos) & (kRing - 1);
#pragma unroll
          for (int k = 0; k < 4; k++)
            if ((SHAPE >> k & 1) && (k == kFirst || cb[k])) ring[ring_field(SHAPE, k) * kRing + e] = kv[k][j];
          ring[kRowField * kRing + e] = base_row + item_row<BLOCK>(j);
        }
        ring_tail += __popcll(b);
        if (ring_tail - ring_head >= kWave) {
          ring_round(kWave);
          ring_head += kWave;
        }
      }
      continue;
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
          if ((P.pmask >> k & 1) && P.out[k + 1]) P.out[k + 1][pos] = kv[k][j];
        pos++;
      }
      continue;
    }
    if (!wave_any(sel)) continue;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (!P.aggr[c]) {
#pragma unroll
        for (int j = 0; j < IPT; j++) av[c][j] = 0;
      } else if (!kEager) {
        load_col<BLOCK, VPT, MODE>(P.cache, P.aggr[c], seg, start, vec, n, lo, sel, av[c]);
      }
    }

    if (SINK == AGGR) {
#pragma unroll
      for (int j = 0; j < IPT; j++)
        if (sel >> j & 1u) acc += (long long)(av[0][j] * av[1][j]);  // int32 product (GPUProcessing.h:1763)
      continue;
    }
    // dense group index (GPUProcessing.h:142) of every selected item first,
    // then the updates, so LDS round trips overlap
    int gh[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++)
      gh[j] = (sel >> j & 1u)
                  ? mod_pos((kv[0][j] - P.minv[0]) * P.uniq[0] + (kv[1][j] - P.minv[1]) * P.uniq[1] +
                                (kv[2][j] - P.minv[2]) * P.uniq[2] + (kv[3][j] - P.minv[3]) * P.uniq[3],
                            P.total_val)
                  : 0;
    if (SINK == GROUPBY_GLOBAL) {
#pragma unroll
      for (int j = 0; j < IPT; j++) {
        if (!(sel >> j & 1u)) continue;
        int2* key = reinterpret_cast<int2*>(P.res + 6 * (size_t)gh[j]);
        key[0] = make_int2(kv[0][j], kv[1][j]);
        key[1] = make_int2(kv[2][j], kv[3][j]);
        atomicAdd(reinterpret_cast<unsigned long long*>(key + 2),
                  (unsigned long long)(long long)(av[0][j] - av[1][j]));
      }
    } else {
      uint32_t first = 0;
#pragma unroll
      for (int j = 0; j < IPT; j++) {
        if (!(sel >> j & 1u)) continue;
        const uint32_t bit = 1u << (gh[j] & 31);
        if (!(atomicOr(&touched[gh[j] >> 5], bit) & bit)) first |= 1u << j;
      }
#pragma unroll
      for (int j = 0; j < IPT; j++) {
        if (!(sel >> j & 1u)) continue;
        atomicAdd(&gsum[gh[j]], (unsigned long long)(long long)(av[0][j] - av[1][j]));
        const int vj[4] = {kv[0][j], kv[1][j], kv[2][j], kv[3][j]};
        if ((first >> j & 1u) && !keys_decode(P, vj)) {
          // keys that do not decode from the index are written by the first
          // item of this workgroup that reaches the group; the reference
          // stores them for every item, with the same values
          int2* key = reinterpret_cast<int2*>(P.res + 6 * (size_t)gh[j]);
          key[0] = make_int2(vj[0], vj[1]);
          key[1] = make_int2(vj[2], vj[3]);
          atomicOr(&stored[gh[j] >> 5], 1u << (gh[j] & 31));
        }
      }
    }
  }

  if constexpr (kCompact) {
    if (ring_tail > ring_head) ring_round(ring_tail - ring_head);
  }
  if (SINK == GROUPBY_LDS) {
    __syncthreads();
    for (int h = threadIdx.x; h < P.total_val; h += BLOCK) {
      if (!(touched[h >> 5] >> (h & 31) & 1u)) continue;
      unsigned long long* sum = reinterpret_cast<unsigned long long*>(P.res + 6 * (size_t)h + 4);
      if (!P.kdecode) {
        atomicAdd(sum, gsum[h]);
        continue;
      }
      // the workgroup that finds the sum still 0 (the first to add, or one
      // after contributions that cancelled) stores the keys decoded from the
      // index, unless this workgroup stored keys that do not decode
      if (atomicAdd(sum, gsum[h]) == 0ull && !(stored[h >> 5] >> (h & 31) & 1u)) {
        int kk[4];
#pragma unroll
        for (int k = 0; k < 4; k++) kk[k] = P.uniq[k] != 0 ? P.minv[k] + (h / P.uniq[k]) % P.krange[k] : 0;
        int2* key = reinterpret_cast<int2*>(P.res + 6 * (size_t)h);
        key[0] = make_int2(kk[0], kk[1]);
        key[1] = make_int2(kk[2], kk[3]);
      }
    }
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
  SegSlots<3> S;  // filter, key, value columns
  const int* const scols[3] = {B.fcol, B.key, B.val};
  const unsigned sused = (B.fcol ? 1u : 0u) | 2u | (B.val ? 4u : 0u);
  int tile0, tile1;
  tile_range(ntiles, tile0, tile1);
  int kmn = B.kmin_init, kmx = B.kmax_init;
  for (int tile = tile0; tile < tile1; tile++) {
    const int start = tile * TILE;
    const int n = min(TILE, B.num_tuples - start);
    const bool full = n == TILE;
    const uint32_t rmask = full ? (IPT == 32 ? 0xffffffffu : (1u << IPT) - 1u)
                                : range_mask<BLOCK, VPT>(n);
    const bool vec = full && B.aligned;
    uint32_t sel = rmask;
    int seg = 0, base_row = 0, row[IPT];
    if (MODE == SEG) {
      seg_slots(S, B.seg_group, B.start_offset, start, scols, sused);
      seg = S.seg;
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
    if constexpr (MODE != OFF) {
      if (B.fcol) {
        load_eager<BLOCK, VPT, MODE>(B.cache, B.fcol, S.slot[0], start, vec, n, x);
        sel &= pred_mode<VPT>(x, B.fmode, B.c1, B.c2, rmask);
      }
      if (!wave_any(sel)) continue;
      load_eager<BLOCK, VPT, MODE>(B.cache, B.key, S.slot[1], start, vec, n, x);
      if (B.val) load_eager<BLOCK, VPT, MODE>(B.cache, B.val, S.slot[2], start, vec, n, v);
    } else {
      if (!wave_any(sel)) continue;
      load_col<BLOCK, VPT, MODE>(B.cache, B.key, seg, start, vec, n, row, sel, x);
      if (B.val) load_col<BLOCK, VPT, MODE>(B.cache, B.val, seg, start, vec, n, row, sel, v);
    }
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      if (!(sel >> j & 1u)) continue;
      const size_t h = (size_t)phash(x[j], B.num_slots, B.key_min);
      if (B.val) atomicCAS(&B.ht[2 * h], 0, v[j]);
      B.ht[2 * h + 1] = row[j] + 1;
    }
    if (B.kmin_out) {
#pragma unroll
      for (int j = 0; j < IPT; j++)
        if (sel >> j & 1u) {
          kmn = min(kmn, x[j]);
          kmx = max(kmx, x[j]);
        }
    }
  }
  if (B.kmin_out) {
    // wave reduction, then one atomic pair per wave (min and max are
    // associative, so the merge order does not change the result)
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
      kmn = min(kmn, __shfl_xor(kmn, d, kWave));
      kmx = max(kmx, __shfl_xor(kmx, d, kWave));
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
      atomicMin(B.kmin_out, kmn);
      atomicMax(B.kmax_out, kmx);
    }
  }
}


// ------------------------------------------------------------ host helpers


#ifndef MORDRED_BLOCK
#define MORDRED_BLOCK 256
#endif
#ifndef MORDRED_VPT
#define MORDRED_VPT 2
#endif
#ifndef MORDRED_AGGR_VPT
#define MORDRED_AGGR_VPT 1
#endif
#ifndef MORDRED_AGGR_BLOCK
#define MORDRED_AGGR_BLOCK 1024
#endif
#ifndef MORDRED_LDS_BLOCK
#define MORDRED_LDS_BLOCK 1024
#endif
constexpr int kBlock = MORDRED_BLOCK;
constexpr int kVpt = MORDRED_VPT;
constexpr int kLdsBlock = MORDRED_LDS_BLOCK;
// scalar aggregates end with one same-address atomic per workgroup, so they
// use fewer, larger workgroups
constexpr int kAggrBlock = MORDRED_AGGR_BLOCK;
constexpr int kAggrVpt = MORDRED_AGGR_VPT;
constexpr size_t kLdsBudget = 80 * 1024;   // dynamic LDS of one workgroup (2 per CU)
// The 1024-thread group-by kernel runs one workgroup per CU (its registers
// allow no second), so it may use the whole 160 KiB less the static arrays.
constexpr size_t kLdsBudgetOneGroup = 159 * 1024;
constexpr int kCompactMinLen = 1 << 17;    // hash tables above 1 MB get a compact table

// Workgroups the device keeps resident for kernel f (kernels.hip, cached).
int resident_groups(const void* f, int block, size_t lds);

// Persistent grid for ntiles tiles: every workgroup walks the same number of
// tiles (ceil(ntiles / resident)), so a short input does not end with half the
// workgroups running one extra tile.
inline int balanced_grid(int ntiles, int resident) {
  if (ntiles <= 0) return 1;
  resident = std::max(1, resident);
  const int per = (ntiles + resident - 1) / resident;
  return (ntiles + per - 1) / per;
}

inline size_t group_lds_bytes(int sink, int total_val) {
  if (sink != GROUPBY_LDS) return 0;
  return ((size_t)total_val + 2 * (((size_t)total_val + 63) / 64)) * sizeof(unsigned long long);
}

// 32-bit words of a compact table of len entries at width w (entries never
// straddle words).
inline size_t staged_words(int len, int w) {
  const int per = 32 / w;
  return ((size_t)len + per - 1) / per;
}

inline size_t ring_lds_bytes(int block, int shape) {
  return (size_t)(block / kWave) * kRing * ring_fields(shape) * sizeof(int);
}

inline size_t pipe_lds_bytes(const Pipe& P, int sink) {
  size_t b = group_lds_bytes(sink, P.total_val);
  b += (((size_t)P.lds_words + 1) / 2) * sizeof(unsigned long long);
  return b;
}

template <int BLOCK, int VPT, int MODE, int SINK, int SHAPE = kShapeGeneric>
int launch_pipe(const Pipe& Pin, hipStream_t stream) {
  constexpr int TILE = BLOCK * 4 * VPT;
  Pipe P = Pin;
  size_t lds = pipe_lds_bytes(P, SINK);
  if (shape_compacts(MODE, SINK, SHAPE)) {
    P.ring_off = (int)(lds / sizeof(unsigned long long));
    lds += ring_lds_bytes(BLOCK, SHAPE);
  }
  const void* f = reinterpret_cast<const void*>(&pipeline_kernel<BLOCK, VPT, MODE, SINK, SHAPE>);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const int grid = balanced_grid((P.num_tuples + TILE - 1) / TILE, resident_groups(f, BLOCK, lds));
  pipeline_kernel<BLOCK, VPT, MODE, SINK, SHAPE><<<grid, BLOCK, lds, stream>>>(P);
  return hipGetLastError();
}

// Launchers of the compile-time specialised shapes (spec_*.hip).  They
// return -1 when `shape` has no specialisation, so the caller uses the
// generic kernel.
int launch_groupby_lds_spec(const Pipe& P, int shape, hipStream_t stream);
int launch_aggr_spec(const Pipe& P, int shape, hipStream_t stream);

}  // namespace mordred
