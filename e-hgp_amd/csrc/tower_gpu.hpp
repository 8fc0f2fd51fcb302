#pragma once

#include <cstdint>
#include <vector>

#include "context.hpp"

namespace mhgp {

// Saddles of one order with their resolved arm carriers (device results).
// cls, level, arm_base and carrier point into the context's page-locked
// buffer of the order and stay valid until the next tower call.
struct TowerOrderData {
  uint32_t minima = 0;
  uint32_t saddles = 0;
  uint32_t arms = 0;
  const uint32_t *cls = nullptr;       // exact level class per saddle (canonical order)
  const double *level = nullptr;       // FP64 level per saddle
  const uint32_t *arm_base = nullptr;  // saddles + 1 offsets into carrier
  uint32_t *carrier = nullptr;         // minimum index per arm, kInvalid = host
  std::vector<uint32_t> sup;            // 4 support ids per saddle, only when the host must descend
  std::vector<uint32_t> minima_labels;  // only when the host must descend
  std::vector<uint32_t> saddle_labels;  // 11 per saddle, only when the host must descend
  int64_t steps = 0;
};

void tower_order_device(Context &c, int k, TowerOrderData &out);

}  // namespace mhgp
