// Wave-aggregated append into a device array: one 64-bit atomic per wave
// (the ballot of the lanes that keep a record), each lane's slot from the
// popcount of the kept lanes below it.  Used to compact the undecided cells
// of one level into the next level's input.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccd {

// Slot of this lane's record if `keep`; every active lane of the wave must
// reach the call.  The counter counts every kept record, also past the
// capacity of the array it indexes, so the caller can detect overflow.
__device__ __forceinline__ uint64_t wave_append(bool keep, unsigned long long* counter) {
  const uint64_t m = __ballot(keep);
  if (m == 0) return ~0ull;
  const int leader = __ffsll((unsigned long long)m) - 1;
  const int lane = __lane_id();
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  const uint64_t below = m & ((1ull << lane) - 1ull);
  return keep ? base + (uint64_t)__popcll(below) : ~0ull;
}

}  // namespace ccd
