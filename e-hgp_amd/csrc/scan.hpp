#pragma once

#include "common.hpp"

namespace mhgp {

// In-place inclusive prefix sum of n uint32 values (three-phase, any n).
struct Scanner {
  DevBuf<uint32_t> sums;
  DevBuf<uint32_t> sums2;
  void inclusive(uint32_t *data, int64_t n, hipStream_t s);
};

}  // namespace mhgp
