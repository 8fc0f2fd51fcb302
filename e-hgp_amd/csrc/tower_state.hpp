// Host state of the hgp_reduced tower.
#pragma once

#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "stages.hpp"

namespace mhgp {

// Merge forest of one order; node ids and children are local to the order.
struct Forest {
  std::vector<int32_t> order, kind;
  std::vector<int64_t> level_class;
  std::vector<double> level;
  std::vector<int64_t> child_offset{0};
  std::vector<int64_t> children;
  // Vertical maps.  Contraction elements are the order's minima (points at
  // order one, dense representatives otherwise): per element, the node that
  // first owned its class and that node's level class there; per node, an
  // element of its class at its own level; element -> raw minimum index
  // (empty: identity); per node, its vertical target in the order below
  // (node id local to that order, -1 at order one).
  std::vector<int32_t> acq_node;
  std::vector<int64_t> acq_cls;
  std::vector<int32_t> rep_elem;
  std::vector<uint32_t> raw_of_dense;
  std::vector<int64_t> vtarget;

  int64_t add(int32_t o, int32_t k, int64_t cls, double lev, const int64_t *kids = nullptr, size_t nk = 0) {
    rep_elem.push_back(-1);
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
    rep_elem.reserve(nodes);
    order.reserve(nodes);
    kind.reserve(nodes);
    level_class.reserve(nodes);
    level.reserve(nodes);
    child_offset.reserve(nodes + 1);
    children.reserve(2 * nodes);
  }
};

struct TowerState {
  bool valid = false;
  std::vector<Forest> forests;  // order k at index k - 1
  int64_t node_count = 0, child_count = 0;
  int64_t descent_steps = 0;
  int32_t orders = 0;
  int32_t status = 0;
  int level_exp = 0;  // levels are reported times 2^level_exp (device unit -> input unit)

  // The tower as one node list: orders ascending, node ids and child offsets
  // shifted by the nodes and children of the orders below (one thread per
  // order, each writes its own ranges).
  void copy_out(int32_t *order, int32_t *kind, int64_t *level_class, double *level, int64_t *child_offset,
                int64_t *children) const {
    child_offset[0] = 0;
    std::vector<std::thread> th;
    int64_t nb = 0, cb = 0;
    for (const Forest &F : forests) {
      const int lexp = level_exp;
      th.emplace_back([&F, nb, cb, lexp, order, kind, level_class, level, child_offset, children] {
        const size_t m = F.order.size();
        std::copy(F.order.begin(), F.order.end(), order + nb);
        std::copy(F.kind.begin(), F.kind.end(), kind + nb);
        std::copy(F.level_class.begin(), F.level_class.end(), level_class + nb);
        for (size_t i = 0; i < m; ++i) level[nb + (int64_t)i] = std::ldexp(F.level[i], lexp);
        for (size_t i = 1; i <= m; ++i) child_offset[nb + (int64_t)i] = cb + F.child_offset[i];
        for (size_t j = 0; j < F.children.size(); ++j) children[cb + (int64_t)j] = nb + F.children[j];
      });
      nb += (int64_t)F.order.size();
      cb += (int64_t)F.children.size();
    }
    for (auto &t : th) t.join();
  }

  // Vertical targets as global node ids (orders ascending), -1 at order one.
  void copy_vertical(int64_t *target) const {
    int64_t nb = 0, below = 0;
    for (const Forest &F : forests) {
      for (size_t i = 0; i < F.order.size(); ++i)
        target[nb + (int64_t)i] = (F.vtarget.empty() || F.vtarget[i] < 0) ? -1 : below + F.vtarget[i];
      below = nb;
      nb += (int64_t)F.order.size();
    }
  }
};

}  // namespace mhgp
