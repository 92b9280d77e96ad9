// Compile-time specialised segment-mode group-by kernels (LDS group array):
// one instantiation per set of probed tables, without fact filters, with the
// first BETWEEN filter only, or with both (the fused shapes of the SSB Q2-Q4
// families and of star queries with one fact predicate).
#include "pipeline.hpp"

namespace mordred {

template <int S>
static int go(const Pipe& P, hipStream_t s) {
  return launch_pipe<kLdsBlock, kVpt, SEG, GROUPBY_LDS, S>(P, s);
}

int launch_groupby_lds_spec(const Pipe& P, int shape, hipStream_t stream) {
  switch (shape) {
#define MORDRED_CASE(S) \
  case S: return go<S>(P, stream);
#define MORDRED_CASES(F)                                                              \
  MORDRED_CASE(F | 1) MORDRED_CASE(F | 2) MORDRED_CASE(F | 3) MORDRED_CASE(F | 4)     \
  MORDRED_CASE(F | 5) MORDRED_CASE(F | 6) MORDRED_CASE(F | 7) MORDRED_CASE(F | 8)     \
  MORDRED_CASE(F | 9) MORDRED_CASE(F | 10) MORDRED_CASE(F | 11) MORDRED_CASE(F | 12)  \
  MORDRED_CASE(F | 13) MORDRED_CASE(F | 14) MORDRED_CASE(F | 15)
    MORDRED_CASES(0)
    MORDRED_CASES(0x10)
    MORDRED_CASES(0x30)
#undef MORDRED_CASES
#undef MORDRED_CASE
    default: return -1;
  }
}

}  // namespace mordred
