#pragma once

#include <cstdint>
#include <vector>

#include "context.hpp"

namespace mhgp {

// Saddles of one order with their resolved arm carriers (device results).
struct TowerOrderData {
  uint32_t minima = 0;
  uint32_t saddles = 0;
  std::vector<uint32_t> cls;       // exact level class per saddle (canonical order)
  std::vector<double> level;       // FP64 level per saddle
  std::vector<uint32_t> sup;       // 4 support ids per saddle
  std::vector<uint32_t> arm_base;  // saddles + 1 offsets into carrier
  std::vector<uint32_t> carrier;   // minimum index per arm, kInvalid = host
  std::vector<uint32_t> minima_labels;  // only when the host must descend
  std::vector<uint32_t> saddle_labels;  // 11 per saddle, only when the host must descend
  int64_t steps = 0;
};

void tower_order_device(Context &c, int k, TowerOrderData &out);

}  // namespace mhgp
