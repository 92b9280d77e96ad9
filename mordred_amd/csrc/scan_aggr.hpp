// Segment-mode scalar aggregate scan: filter_probe_aggr_GPU2 and its
// probe-only, filter-only and plain forms (GPUProcessing.h:1693-1775), the
// SSB Q1.x shape.  Per fact row: up to two BETWEEN filters, an optional probe
// of the date table (table 4), and SUM of the int32 product of two columns
// into one int64.
//
// The scan reads 12-20 bytes per row and a workgroup walks only a few tiles
// (a 6M-row fact table is about three 4,096-row tiles per workgroup), so the
// time is set by how early every tile's column loads are in flight:
//   * the next tile's columns are loaded while the current tile is filtered,
//     probed and summed (two tiles in flight per wave);
//   * the hash-table gathers of the current tile are issued before the next
//     tile's loads, so waiting for them (loads complete in issue order) does
//     not wait for the next tile;
//   * when the second aggregate column is the first filter column (Q1.x sums
//     extendedprice * discount and filters on discount) the product reuses the
//     filter's registers instead of loading the column twice;
//   * every load of a column takes one code path (no run-time branch around
//     it): where two paths write the same registers the compiler waits for
//     every load in flight before either.
#pragma once

#include <type_traits>

#include "pipeline.hpp"

namespace mordred {

template <int IPT>
struct AggrTile {
  int f0[IPT], f1[IPT], k[IPT], a0[IPT], a1[IPT];
};

// This thread's 4*VPT items of a tile of cached segment slot `slot`.  VEC
// (a 16-byte aligned cache): one 128-bit load per 4 rows, also for the last,
// partial tile -- every cache slot holds a whole segment of 2^20 ints
// (CacheManager.cu:44 allocates whole segments) and a tile never straddles a
// segment, so the loads stay inside the slot, and the range mask drops the
// items past the end.  One load path per column, with no branch on the tile's
// length: a branch between two load forms that write the same registers makes
// the compiler wait for every load in flight before either form.
template <int BLOCK, int VPT, bool VEC>
__device__ __forceinline__ void load_slot(const int* cache, int slot, int start, int n, int (&x)[4 * VPT]) {
  const int* p = cache + ((size_t)slot << kSegLog) + (start & kSegMask);
  if (VEC) {
#pragma unroll
    for (int v = 0; v < VPT; v++) {
      const v4i t = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(p) + v * BLOCK + (int)threadIdx.x);
      x[4 * v + 0] = t.x; x[4 * v + 1] = t.y; x[4 * v + 2] = t.z; x[4 * v + 3] = t.w;
    }
  } else {
    load_tile<BLOCK, VPT>(p, false, n, x);
  }
}

// Column slots: 0-1 filters, 2 date key, 3-4 aggregates.  A1F0: aggregate
// column 2 is filter column 1 and is not loaded again.
template <int BLOCK, int VPT, int SHAPE, bool VEC, bool A1F0>
__device__ __forceinline__ void aggr_load(const Pipe& P, SegSlots<5>& S, const int* const (&cols)[5],
                                          unsigned used, int start, int n, AggrTile<4 * VPT>& T) {
  seg_slots(S, P.seg_group, P.start_offset, start, cols, used);
  if (SHAPE >> 4 & 1) load_slot<BLOCK, VPT, VEC>(P.cache, S.slot[0], start, n, T.f0);
  if (SHAPE >> 5 & 1) load_slot<BLOCK, VPT, VEC>(P.cache, S.slot[1], start, n, T.f1);
  if (SHAPE >> 3 & 1) load_slot<BLOCK, VPT, VEC>(P.cache, S.slot[2], start, n, T.k);
  load_slot<BLOCK, VPT, VEC>(P.cache, S.slot[3], start, n, T.a0);
  if (!A1F0) load_slot<BLOCK, VPT, VEC>(P.cache, S.slot[4], start, n, T.a1);
}

// SHAPE: bit 3 = date table probed, bit 4 / bit 5 = filter column 1 / 2.
// CT: the host derived a compact date table (its width is read on the
// device).  Without one the probe reads the hash table only, with no branch
// whose two sides the gathered values would have to be merged from (such a
// merge makes the compiler wait for the gathers before the next tile's loads).
// VEC: the cache is 16-byte aligned (load_slot).  Both aggregate columns
// are present (the host launches nothing when one is absent: every product
// would be 0).
template <int BLOCK, int VPT, int SHAPE, bool CT, bool VEC, bool A1F0>
__global__ __launch_bounds__(BLOCK) void aggr_scan_kernel(Pipe P) {
  static_assert((SHAPE & 0x7) == 0, "aggregates probe the date table only");
  constexpr int IPT = 4 * VPT;
  constexpr int TILE = BLOCK * IPT;
  constexpr bool kProbe = SHAPE >> 3 & 1;
  __shared__ long long red[BLOCK / kWave];
  extern __shared__ uint32_t ldsw[];

  // compact date table (probe_tables.hpp), if the host derived one: entry
  // width, and whether its packed words are staged in LDS
  int cw = 0, csh = 0;
  bool stg = false;
  if (CT && kProbe && P.chdr[3]) {
    const CompactHdr h = *P.chdr[3];
    cw = h.w;
    csh = h.shift;
    if ((P.lds_stage >> 3 & 1) && cw > 0) {
      const int nwords = (P.len[3] + (1 << csh) - 1) >> csh;
      if (nwords <= P.lds_words) {
        stg = true;
        for (int i = threadIdx.x; i < nwords; i += BLOCK) ldsw[i] = P.cwords[3][i];
        __syncthreads();
      }
    }
  }

  const int* const cols[5] = {P.fcol[0], P.fcol[1], P.key[3], P.aggr[0], P.aggr[1]};
  constexpr unsigned used = ((SHAPE >> 4 & 1) ? 1u : 0u) | ((SHAPE >> 5 & 1) ? 2u : 0u) | (kProbe ? 4u : 0u) | 8u |
                            (A1F0 ? 0u : 16u);
  const int ntiles = (P.num_tuples + TILE - 1) / TILE;
  int t0, t1;
  tile_range(ntiles, t0, t1);
  long long acc = 0;
  SegSlots<5> S;
  AggrTile<IPT> cur, nxt;
  // One tile: filters, probe gathers, then (PF) the next tile's loads, then
  // the probe results and the sum.  The last tile of the range is a separate
  // copy without the next-tile loads, so that no branch around them merges
  // two counts of loads in flight (the compiler would then wait for all of
  // them before using the gathers).
  auto step = [&](auto pf, int t) {
    constexpr bool PF = decltype(pf)::value;
    const int start = t * TILE;
    const int n = min(TILE, P.num_tuples - start);
    uint32_t sel = n == TILE ? (IPT == 32 ? 0xffffffffu : (1u << IPT) - 1u) : range_mask<BLOCK, VPT>(n);
    if (SHAPE >> 4 & 1) sel &= pred_between<VPT>(cur.f0, P.cmp[0], P.cmp[1]);
    if (SHAPE >> 5 & 1) sel &= pred_between<VPT>(cur.f1, P.cmp[2], P.cmp[3]);

    // issue this tile's probe gathers: every item loads (unselected items
    // read slot 0, one shared cache line), so that no branch sits between
    // the gathers and their use and the compiler keeps them in flight
    int hs[IPT];
    unsigned long long sv[IPT];
    uint32_t wv[IPT];
    if (kProbe) {
#pragma unroll
      for (int j = 0; j < IPT; j++) hs[j] = (sel >> j & 1u) ? phash(cur.k[j], P.len[3], P.minkey[3]) : 0;
      if (CT && cw > 0) {
        if (stg) {
#pragma unroll
          for (int j = 0; j < IPT; j++) wv[j] = ldsw[hs[j] >> csh];
        } else {
          const uint32_t* cwk = P.cwords[3];
#pragma unroll
          for (int j = 0; j < IPT; j++) wv[j] = cwk[hs[j] >> csh];
        }
      } else {
        const int* ht = P.ht[3];
#pragma unroll
        for (int j = 0; j < IPT; j++)
          sv[j] = *reinterpret_cast<const unsigned long long*>(ht + 2 * (size_t)hs[j]);
      }
    }
    if constexpr (PF) {
      // keep the gathers ahead of the next tile's loads
      __builtin_amdgcn_sched_barrier(0);
      aggr_load<BLOCK, VPT, SHAPE, VEC, A1F0>(P, S, cols, used, start + TILE, min(TILE, P.num_tuples - start - TILE),
                                             nxt);
      __builtin_amdgcn_sched_barrier(0);
    }

    if (kProbe) {
      if (CT && cw > 0) {
        const uint32_t emask = (1u << cw) - 1u;
#pragma unroll
        for (int j = 0; j < IPT; j++)
          if (((wv[j] >> ((hs[j] & ((1 << csh) - 1)) * cw)) & emask) == 0u) sel &= ~(1u << j);
      } else {
        // BlockProbeGroupByGPU: hit iff the 64-bit slot != 0
#pragma unroll
        for (int j = 0; j < IPT; j++)
          if (sv[j] == 0ull) sel &= ~(1u << j);
      }
    }
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      if (!(sel >> j & 1u)) continue;
      const unsigned a = (unsigned)cur.a0[j], b = (unsigned)(A1F0 ? cur.f0[j] : cur.a1[j]);
      acc += (long long)(int)(a * b);  // int32 product (GPUProcessing.h:1763)
    }
  };
  if (t0 < t1) {
    aggr_load<BLOCK, VPT, SHAPE, VEC, A1F0>(P, S, cols, used, t0 * TILE, min(TILE, P.num_tuples - t0 * TILE), cur);
    int t = t0;
    for (; t + 1 < t1; t++) {
      step(std::true_type{}, t);
      cur = nxt;
    }
    step(std::false_type{}, t);
  }
  const long long s = block_sum<BLOCK>(acc, red);
  if (threadIdx.x == 0 && s != 0) atomicAdd(reinterpret_cast<unsigned long long*>(P.res + 4), (unsigned long long)s);
}

template <int BLOCK, int VPT, int SHAPE, bool CT, bool VEC, bool A1F0>
int launch_aggr_scan_as(const Pipe& P, hipStream_t stream) {
  constexpr int TILE = BLOCK * 4 * VPT;
  const size_t lds = CT && (P.lds_stage >> 3 & 1) ? (size_t)P.lds_words * sizeof(uint32_t) : 0;
  const void* f = reinterpret_cast<const void*>(&aggr_scan_kernel<BLOCK, VPT, SHAPE, CT, VEC, A1F0>);
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const int grid = balanced_grid((P.num_tuples + TILE - 1) / TILE, resident_groups(f, BLOCK, lds));
  aggr_scan_kernel<BLOCK, VPT, SHAPE, CT, VEC, A1F0><<<grid, BLOCK, lds, stream>>>(P);
  return hipGetLastError();
}

template <int BLOCK, int VPT, int SHAPE, bool A1F0>
int launch_aggr_scan_dup(const Pipe& P, hipStream_t stream) {
  if (!P.aligned) return launch_aggr_scan_as<BLOCK, VPT, SHAPE, true, false, A1F0>(P, stream);
  if ((SHAPE >> 3 & 1) && P.chdr[3]) return launch_aggr_scan_as<BLOCK, VPT, SHAPE, true, true, A1F0>(P, stream);
  return launch_aggr_scan_as<BLOCK, VPT, SHAPE, false, true, A1F0>(P, stream);
}

// SUM(a1 * a2) of the selected rows.  The int32 product commutes, so the
// aggregate that is also filter column 1 (if any) is taken as the second.
template <int BLOCK, int VPT, int SHAPE>
int launch_aggr_scan(const Pipe& Pin, hipStream_t stream) {
  Pipe P = Pin;
  if (!P.aggr[0] || !P.aggr[1]) return hipSuccess;  // every product is 0
  if constexpr ((SHAPE >> 4 & 1) != 0) {
    if (P.aggr[0] == P.fcol[0] && P.aggr[1] != P.fcol[0]) std::swap(P.aggr[0], P.aggr[1]);
    if (P.aggr[1] == P.fcol[0]) return launch_aggr_scan_dup<BLOCK, VPT, SHAPE, true>(P, stream);
  }
  return launch_aggr_scan_dup<BLOCK, VPT, SHAPE, false>(P, stream);
}

}  // namespace mordred
