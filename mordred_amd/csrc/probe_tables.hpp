// Compact probe tables: an exact, smaller stand-in for a perfect-hash table,
// derived on the device right before a fact kernel probes it.
//
// A probe of slot h of a reference hash table ([value, row+1] int32 pairs,
// BlockLibrary.cuh:613-614) yields (hit, result):
//   value probe (BlockProbeGroupByGPU):  hit = 64-bit slot != 0, result = value word
//   row probe   (BlockProbeGPU):         hit = row word != 0,    result = row word - 1
// With vmin the smallest result over hit slots, the entry
//   e(h) = hit ? result - vmin + 1 : 0
// carries the same information in ceil(log2(range+2)) bits.  Entries are packed
// at a power-of-two width w in {1,2,4,8,16} (32/w per 32-bit word, never split
// across words); a table whose range needs more than 16 bits is not compacted
// (w = 0) and the kernel probes the hash table itself.  A table built with no
// value column (e.g. a filtered supplier table of a query that does not group
// by it) has a constant result and compacts to one bit per slot.
//
// The caller's hash tables are only read.  Two launches per query: `stats`
// writes a (min, max, hits) partial per workgroup, `pack` reduces the partials
// (every workgroup redundantly, so no third launch or pre-zeroed memory is
// needed), writes the header and the packed words.
//
// A large table also gets an existence bitmap (one bit per slot).  When few
// slots hit, a probe reads the bitmap first (len/8 bytes, which stays in L2
// while the fact columns stream through) and fetches the wider entry only on
// a hit, instead of pulling a cache line of entries from MALL for every probe.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tile.hpp"

namespace mordred {

struct CompactHdr {
  int w;      // entry width in bits; 0 = not compacted
  int shift;  // log2(entries per 32-bit word)
  int vmin;
  int any;    // number of hit slots
};

constexpr int kPrepMaxTables = 4;
#ifndef MORDRED_STATS_GROUPS
#define MORDRED_STATS_GROUPS 512
#endif
#ifndef MORDRED_PACK_GROUPS
#define MORDRED_PACK_GROUPS 1024
#endif
constexpr int kStatsGroups = MORDRED_STATS_GROUPS;  // workgroups of the stats pass per table
constexpr int kPackGroups = MORDRED_PACK_GROUPS;    // most workgroups of the pack pass per table
constexpr int kPrepBlock = 256;

struct Prep {
  const int* ht[kPrepMaxTables];
  int len[kPrepMaxTables];
  int row_probe;                    // 1: row probes, 0: value probes
  int4* partial[kPrepMaxTables];    // kStatsGroups x (min, max, any, -)
  CompactHdr* hdr[kPrepMaxTables];
  uint32_t* words[kPrepMaxTables];
  uint32_t* bits[kPrepMaxTables];   // existence bitmap, or nullptr
};

__device__ __forceinline__ bool slot_decode(int2 v, int row_probe, int& r) {
  if (row_probe) {
    r = v.y - 1;
    return v.y != 0;
  }
  r = v.x;
  return (v.x | v.y) != 0;
}

__device__ __forceinline__ int2 slot_load(const int* ht, int s) {
  return *reinterpret_cast<const int2*>(ht + 2 * (size_t)s);
}

__device__ __forceinline__ bool slot_result(const int* ht, int s, int row_probe, int& r) {
  return slot_decode(slot_load(ht, s), row_probe, r);
}

// The two pre-pass kernels are defined in exactly one translation unit
// (kernels.hip, which defines MORDRED_DEFINE_PREP_KERNELS).
__global__ void prep_stats_kernel(Prep P);
__global__ void prep_pack_kernel(Prep P);

#ifdef MORDRED_DEFINE_PREP_KERNELS
// blockIdx.y = table, blockIdx.x < kStatsGroups
__global__ __launch_bounds__(kPrepBlock) void prep_stats_kernel(Prep P) {
  const int t = blockIdx.y;
  const int* ht = P.ht[t];
  const int len = P.len[t];
  int mn = 0x7fffffff, mx = (int)0x80000000, any = 0;
  constexpr int kStride = kStatsGroups * kPrepBlock;
  int s = blockIdx.x * kPrepBlock + threadIdx.x;
  // four independent slot loads in flight per thread, then the remainder
  for (; s + 3 * kStride < len; s += 4 * kStride) {
    int2 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = slot_load(ht, s + u * kStride);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int r;
      if (slot_decode(v[u], P.row_probe, r)) {
        mn = min(mn, r);
        mx = max(mx, r);
        any++;
      }
    }
  }
  for (; s < len; s += kStride) {
    int r;
    if (slot_result(ht, s, P.row_probe, r)) {
      mn = min(mn, r);
      mx = max(mx, r);
      any++;
    }
  }
  __shared__ int smn[kPrepBlock / kWave], smx[kPrepBlock / kWave], sany[kPrepBlock / kWave];
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    mn = min(mn, __shfl_xor(mn, d, kWave));
    mx = max(mx, __shfl_xor(mx, d, kWave));
    any += __shfl_xor(any, d, kWave);
  }
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0) { smn[wave] = mn; smx[wave] = mx; sany[wave] = any; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kPrepBlock / kWave; w++) {
      mn = min(mn, smn[w]);
      mx = max(mx, smx[w]);
      any += sany[w];
    }
    P.partial[t][blockIdx.x] = make_int4(mn, mx, any, 0);
  }
}

// Entry width for results in [vmin, vmax]: entries take values 0..range+1.
__device__ __forceinline__ void choose_width(int vmin, int vmax, int any, int& w, int& shift) {
  if (!any) { w = 1; shift = 5; return; }
  const unsigned long long top = (unsigned long long)((long long)vmax - (long long)vmin) + 1ull;
  int bits = 0;
  while (bits < 64 && (top >> bits) != 0ull) bits++;
  if (bits <= 1) { w = 1; shift = 5; }
  else if (bits <= 2) { w = 2; shift = 4; }
  else if (bits <= 4) { w = 4; shift = 3; }
  else if (bits <= 8) { w = 8; shift = 2; }
  else if (bits <= 16) { w = 16; shift = 1; }
  else { w = 0; shift = 0; }
}

// blockIdx.y = table; grid-stride over slots
__global__ __launch_bounds__(kPrepBlock) void prep_pack_kernel(Prep P) {
  const int t = blockIdx.y;
  const int len = P.len[t];
  __shared__ int4 red[kPrepBlock / kWave];
  int mn = 0x7fffffff, mx = (int)0x80000000, any = 0;
  for (int i = threadIdx.x; i < kStatsGroups; i += kPrepBlock) {
    const int4 p = P.partial[t][i];
    mn = min(mn, p.x);
    mx = max(mx, p.y);
    any += p.z;
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    mn = min(mn, __shfl_xor(mn, d, kWave));
    mx = max(mx, __shfl_xor(mx, d, kWave));
    any += __shfl_xor(any, d, kWave);
  }
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0) red[wave] = make_int4(mn, mx, any, 0);
  __syncthreads();
  mn = 0x7fffffff;
  mx = (int)0x80000000;
  any = 0;
  for (int i = 0; i < kPrepBlock / kWave; i++) {
    mn = min(mn, red[i].x);
    mx = max(mx, red[i].y);
    any += red[i].z;
  }
  int w, shift;
  choose_width(mn, mx, any, w, shift);
  const int vmin = any ? mn : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CompactHdr h;
    h.w = w; h.shift = shift; h.vmin = vmin; h.any = any;
    *P.hdr[t] = h;
  }
  // One slot per lane, one coalesced pass: the bitmap words of a wave's 64
  // slots come from one ballot, a packed word from an OR over the 2^shift
  // consecutive lanes that hold its entries (groups never straddle a wave).
  const int* ht = P.ht[t];
  uint32_t* bits = P.bits[t];
  uint32_t* words = w ? P.words[t] : nullptr;
  const int nb = (len + 31) >> 5;
  const int epw = 1 << shift;
  const int stride = gridDim.x * kPrepBlock;
  // the slots of up to four passes are loaded before any is packed
  for (int base0 = blockIdx.x * kPrepBlock; base0 < len; base0 += 4 * stride) {
    int2 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int s = base0 + u * stride + (int)threadIdx.x;
      v[u] = s < len ? slot_load(ht, s) : make_int2(0, 0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int base = base0 + u * stride;
      if (base >= len) break;
      const int s = base + (int)threadIdx.x;
      int r = 0;
      const bool hit = s < len && slot_decode(v[u], P.row_probe, r);
      if (bits) {
        const unsigned long long b = __ballot(hit);
        if (lane == 0) {
          const int wb = s >> 5;
          if (wb < nb) bits[wb] = (uint32_t)b;
          if (wb + 1 < nb) bits[wb + 1] = (uint32_t)(b >> 32);
        }
      }
      if (words) {
        uint32_t c = hit ? (uint32_t)(r - vmin + 1) << ((s & (epw - 1)) * w) : 0u;
        for (int d = 1; d < epw; d <<= 1) c |= __shfl_xor(c, d, kWave);
        if ((s & (epw - 1)) == 0 && s < len) words[s >> shift] = c;
      }
    }
  }
}

#endif  // MORDRED_DEFINE_PREP_KERNELS

// Bytes of the packed words at the widest compact width (16 bits per slot).
inline size_t compact_words_bytes(int len) { return (((size_t)len + 1) / 2) * sizeof(uint32_t); }
// Bytes of an existence bitmap.
inline size_t compact_bits_bytes(int len) { return (((size_t)len + 31) / 32) * sizeof(uint32_t); }

}  // namespace mordred
