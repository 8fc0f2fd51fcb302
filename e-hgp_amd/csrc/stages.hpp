// Device buffers of the stages after the LBVH build.
#pragma once

#include <vector>

#include "../host/exact.hpp"
#include "context.hpp"

namespace mhgp {

struct HostMirror {
  std::vector<double> xyz;  // canonical coordinates, interleaved
  ExactCloud exact;
  bool exact_ready = false;
};

struct EmstBuffers {
  DevBuf<uint32_t> label, next_label;
  DevBuf<uint32_t> node_label;     // uniform component of an internal node, or kInvalid
  DevBuf<uint32_t> best_v;         // per point: best external neighbour
  DevBuf<double> best_d;           // its FP64 squared distance
  DevBuf<uint8_t> best_flag;       // 1: near tie not settled on the device
  DevBuf<uint64_t> comp_d;         // per component: min squared distance bits
  DevBuf<uint64_t> comp_e;         // per component: min (u, v) among equal distances
  DevBuf<uint32_t> comp_amb;       // per component: ambiguity flag
  DevBuf<uint32_t> comp_cw, comp_ce; // entries in the rounding window / exact equal
  DevBuf<uint32_t> edges;          // accepted edges (u, v)
  DevBuf<uint32_t> counters;
  DevBuf<uint64_t> kappa_key;      // FP64 squared length bits of each edge
  DevBuf<uint32_t> kappa_idx;      // edge index, sorted with the keys
  DevBuf<uint32_t> hook;           // per component: the component it hooks to this round
  int64_t bytes() const {
    return (int64_t)(label.bytes() + next_label.bytes() + node_label.bytes() + best_v.bytes() + best_d.bytes() +
                     best_flag.bytes() + comp_d.bytes() + comp_e.bytes() + comp_amb.bytes() + comp_cw.bytes() +
                     comp_ce.bytes() + edges.bytes() + counters.bytes() + kappa_key.bytes() + kappa_idx.bytes() +
                     hook.bytes());
  }
};

struct CatalogBuffers;
struct TowerState;

}  // namespace mhgp
