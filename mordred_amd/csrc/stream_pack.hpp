// Bit-packed fact segments for the PCIe stream.  A non-resident segment that
// the CPU/GPU split streams to the GPU for its prefix operators crosses PCIe
// as values - base in `bits` bits each (LSB first in 32-bit words; a segment
// takes packed_words(bits) words, one of them slack so that a field's second
// word is always readable), and is unpacked on the device into the staging
// slot where the plain copy would have landed.  No counterpart in the
// reference, which copies 4-byte values (CacheManager.cu:306-361).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mordred {

constexpr int kPackSegLog = 20;  // values per segment: the engine's segment size
constexpr int kPackMaxSegs = 16;
constexpr int kPackMaxBits = 24;  // wider values cross PCIe unpacked

__host__ __device__ inline long long packed_words(int bits) { return (((long long)1 << kPackSegLog) * bits + 31) / 32 + 1; }

struct UnpackBatch {
  const uint32_t* packed;  // segment i at packed + i * packed_words(bits)
  int* cache;              // segment i's values to cache + slot[i] << kPackSegLog
  int bits, base, nseg;
  int slot[kPackMaxSegs];
  int rows[kPackMaxSegs];
};

hipError_t unpack_segments(const UnpackBatch& u, hipStream_t s);

}  // namespace mordred
