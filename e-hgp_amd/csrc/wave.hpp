// Wavefront helpers (device code only).
#pragma once

#include "common.hpp"

namespace mhgp {

// Slot of this lane in a global append list, one atomic per wave: lanes with
// pred true get consecutive slots in lane order; the others get no slot.
// Every lane of the wave that reaches the call must make it.
__device__ inline unsigned long long wave_append(bool pred, unsigned long long *counter) {
  const uint64_t mask = __ballot(pred);
  if (!mask) return 0;
  const int lane = (int)(threadIdx.x & 63u);
  const int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned long long at = 0;
  if (lane == leader) at = atomicAdd(counter, (unsigned long long)__popcll(mask));
  at = __shfl(at, leader);
  return at + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
}

}  // namespace mhgp
