#pragma once

#include "common.hpp"
#include "scan.hpp"

namespace mhgp {

// Stable LSD radix sort of (key, value) pairs over bits [begin_bit, end_bit).
struct RadixSorter {
  DevBuf<uint64_t> tmp_keys;
  DevBuf<uint32_t> tmp_vals;
  DevBuf<uint32_t> counts;
  DevBuf<uint32_t> flags;
  Scanner scanner;
  void sort(uint64_t *keys, uint32_t *vals, int64_t n, int begin_bit, int end_bit,
            hipStream_t stream);
};

}  // namespace mhgp
