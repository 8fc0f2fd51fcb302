// Device buffers and host records of the critical-sphere catalogue.
#pragma once

#include <vector>

#include "stages.hpp"

namespace mhgp {

struct CatalogBuffers {
  bool valid = false;
  int k_max = 0;
  int s_max = 0;
};

inline int64_t catalog_device_bytes(const CatalogBuffers &) { return 0; }

}  // namespace mhgp
