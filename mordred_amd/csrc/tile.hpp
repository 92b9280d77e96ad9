// Wave64 tile primitives for the gfx950 query kernels.
//
// These replace the Crystal block primitives (crystal/load.cuh, pred.cuh,
// reduce.cuh, join.cuh) and the engine's block library
// (src/gpudb/BlockLibrary.cuh) with a CDNA4 layout:
//   * a thread owns VPT groups of 4 CONSECUTIVE rows, so every column read of a
//     full tile is one 128-bit global_load_dwordx4 per group, marked
//     non-temporal so that scans do not evict hash tables from L2/MALL;
//   * selection is a per-thread bit mask (bit j = item j) instead of an int
//     array of flags;
//   * block reductions and the compaction prefix sum use 64-lane wave
//     shuffles, then one LDS exchange between waves.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mordred {

constexpr int kSegLog = 20;                 // src/gpudb/common.h:73
constexpr int kSegSize = 1 << kSegLog;
constexpr int kSegMask = kSegSize - 1;
constexpr int kWave = 64;

typedef int v4i __attribute__((ext_vector_type(4)));

// x % m (C semantics, m > 0) without an integer division when 0 <= x < m,
// the common case for perfect-hash slots and dense group indexes; only lanes
// outside that range take the division.
__device__ __forceinline__ int mod_pos(int x, int m) {
  int r = x;
  if ((unsigned)x >= (unsigned)m) r = x % m;
  return r;
}

// HASH(X,Y,Z) = ((X-Z) % Y)  (crystal/join.cuh:3)
__device__ __forceinline__ int phash(int key, int len, int min_key) { return mod_pos(key - min_key, len); }

// Value of a cached column at fact/dimension row `row`: the segment table maps
// row>>20 to a cache slot (BlockLibrary.cuh:554).  64-bit slot arithmetic.
__device__ __forceinline__ int gather(const int* __restrict__ cache, const int* __restrict__ idx,
                                      int row) {
  return cache[((size_t)idx[row >> kSegLog] << kSegLog) + (row & kSegMask)];
}

// Load this thread's 4*VPT items of a contiguous tile starting at p.
// vec: full tile and p 16-byte aligned -> one dwordx4 per group of 4 rows.
template <int BLOCK, int VPT>
__device__ __forceinline__ void load_tile(const int* __restrict__ p, bool vec, int n,
                                          int (&x)[4 * VPT]) {
  const int tid = threadIdx.x;
  if (vec) {
#pragma unroll
    for (int v = 0; v < VPT; v++) {
      v4i t = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(p) + v * BLOCK + tid);
      x[4 * v + 0] = t.x; x[4 * v + 1] = t.y; x[4 * v + 2] = t.z; x[4 * v + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int v = 0; v < VPT; v++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        int r = (v * BLOCK + tid) * 4 + e;
        x[4 * v + e] = r < n ? p[r] : 0;
      }
  }
}

// Local row index (0..TILE) of item j of this thread.
template <int BLOCK>
__device__ __forceinline__ int item_row(int j) {
  return ((j >> 2) * BLOCK + (int)threadIdx.x) * 4 + (j & 3);
}

// In-range mask of the items of this thread for a tile holding n rows.
template <int BLOCK, int VPT>
__device__ __forceinline__ uint32_t range_mask(int n) {
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < 4 * VPT; j++)
    if (item_row<BLOCK>(j) < n) m |= 1u << j;
  return m;
}

__device__ __forceinline__ bool wave_any(uint32_t m) { return __ballot(m != 0) != 0ull; }

// 64-bit sum over the workgroup; result valid in thread 0.  `scratch` holds
// one value per wave.
template <int BLOCK>
__device__ __forceinline__ long long block_sum(long long x, long long* scratch) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0) scratch[wave] = x;
  __syncthreads();
  long long s = 0;
  if (threadIdx.x == 0)
#pragma unroll
    for (int w = 0; w < BLOCK / kWave; w++) s += scratch[w];
  return s;
}

// Exclusive prefix of `count` over the workgroup plus one global reservation
// of the workgroup total on `*total` (the role of cub::BlockScan + atomicAdd
// in GPUProcessing.h:583-586).  Returns this thread's first output position.
// `scratch` holds BLOCK/64 + 1 ints.  Ends with a barrier so scratch can be
// reused by the caller's next tile.
template <int BLOCK>
__device__ __forceinline__ int block_reserve(int count, int* total, int* scratch) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  int x = count;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  if (lane == kWave - 1) scratch[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / kWave; w++) {
      int t = scratch[w];
      scratch[w] = run;
      run += t;
    }
    scratch[BLOCK / kWave] = run > 0 ? atomicAdd(total, run) : 0;
  }
  __syncthreads();
  int pos = scratch[BLOCK / kWave] + scratch[wave] + x - count;
  __syncthreads();
  return pos;
}

}  // namespace mordred
