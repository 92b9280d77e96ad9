// Compile-time specialised segment-mode scalar aggregates (SUM a1*a2): with
// or without the date probe, with any combination of the two BETWEEN filters
// (the SSB Q1.x shapes), run by the prefetching aggregate scan.
#include "scan_aggr.hpp"

namespace mordred {

template <int S>
static int go(const Pipe& P, hipStream_t s) {
  return launch_aggr_scan<kAggrBlock, kAggrVpt, S>(P, s);
}

int launch_aggr_spec(const Pipe& P, int shape, hipStream_t stream) {
  switch (shape) {
    case 0x00: return go<0x00>(P, stream);
    case 0x08: return go<0x08>(P, stream);
    case 0x10: return go<0x10>(P, stream);
    case 0x18: return go<0x18>(P, stream);
    case 0x20: return go<0x20>(P, stream);
    case 0x28: return go<0x28>(P, stream);
    case 0x30: return go<0x30>(P, stream);
    case 0x38: return go<0x38>(P, stream);
    default: return -1;
  }
}

}  // namespace mordred
