#pragma once

#include "common.hpp"

namespace mhgp {

// In-place inclusive prefix sum of n uint32 values (three-phase, any n).
// Stream-ordered: the level buffers persist between calls, nothing waits.
struct Scanner {
  DevBuf<uint32_t> level[4];
  DevBuf<uint64_t> level64[4];
  void inclusive(uint32_t *data, int64_t n, hipStream_t s, int depth = 0);
  // In-place inclusive prefix maximum of n uint64 values (the bit patterns
  // of non-negative doubles order like the doubles).
  void inclusive_max(uint64_t *data, int64_t n, hipStream_t s, int depth = 0);
};

}  // namespace mhgp
