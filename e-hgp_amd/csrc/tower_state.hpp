// Host state of the hgp_reduced tower.
#pragma once

#include <vector>

#include "stages.hpp"

namespace mhgp {

struct TowerState {
  bool valid = false;
};

}  // namespace mhgp
