// Host state of the hgp_reduced tower.
#pragma once

#include <cstdint>
#include <vector>

#include "stages.hpp"

namespace mhgp {

struct TowerState {
  bool valid = false;
  std::vector<int32_t> order, kind;
  std::vector<int64_t> level_class;
  std::vector<double> level;
  std::vector<int64_t> child_offset{0};
  std::vector<int64_t> children;
  int64_t descent_steps = 0;
  int32_t orders = 0;
  int32_t status = 0;

  int64_t add(int32_t o, int32_t k, int64_t cls, double lev, const int64_t *kids = nullptr, size_t nk = 0) {
    order.push_back(o);
    kind.push_back(k);
    level_class.push_back(cls);
    level.push_back(lev);
    children.insert(children.end(), kids, kids + nk);
    child_offset.push_back((int64_t)children.size());
    return (int64_t)order.size() - 1;
  }
  int64_t add(int32_t o, int32_t k, int64_t cls, double lev, const std::vector<int64_t> &kids) {
    return add(o, k, cls, lev, kids.data(), kids.size());
  }
  void reserve(size_t nodes) {
    order.reserve(nodes);
    kind.reserve(nodes);
    level_class.reserve(nodes);
    level.reserve(nodes);
    child_offset.reserve(nodes + 1);
    children.reserve(2 * nodes);
  }
};

}  // namespace mhgp
