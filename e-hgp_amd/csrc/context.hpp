// Per-cloud device state of the MI355X MorseHGP3D backend.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "../../include/mhgp.h"
#include "radix_sort.hpp"

namespace mhgp {

struct CatalogBuffers;   // supports.hpp
struct EmstBuffers;      // emst.hpp
struct HostMirror;       // pipeline.cpp
struct TowerState;       // tower.cpp
struct TowerWork;        // tower_gpu.hpp
struct AnchorBatch;      // supports.hip (batched first anchor pass)

struct Context {
  int device = 0;
  int64_t capacity_bytes = 0;  // device memory bound of the context (0: none)
  // batched catalogue: the rendezvous and this cloud's slot in it (none: plain)
  AnchorBatch *batch = nullptr;
  int batch_slot = -1;
  // device lanes of the tower (streams with their own workspaces); a server
  // slot uses one, since the server already overlaps whole requests
  int tower_lane_count = 2;
  hipStream_t stream = nullptr;
  bool poisoned = false;
  bool count_flops = false;  // FP64 / FP32 operation counts of the anchor kernel (stats.anchor_flops, anchor_flops32)
  // largest candidate region an anchor's enumeration takes; larger ones stay
  // open and the catalogue is conditional (the enumeration budget)
  uint32_t region_limit = 2048;
  std::string last_error;

  int64_t n = 0;
  bool lbvh_ready = false;
  // Device coordinates (X..Z, LX..LZ, node boxes, bounds) are the canonical
  // ones times 2^scale_exp, an exact power-of-two change of unit chosen at
  // upload so that the cloud's extent is near 1 for clouds far from that
  // scale (their FP32 bounds and FP64 predicates would otherwise leave the
  // normal range).  Signs, orders and ranks do not change; squared levels
  // are divided by 4^scale_exp on their way out.
  int scale_exp = 0;

  // input (source order, AoS) and canonical cloud (PointId order, SoA)
  DevBuf<double> src;
  DevBuf<double> X, Y, Z;
  DevBuf<uint32_t> source_index;  // PointId -> source index
  DevBuf<uint64_t> key0, key1, key2;
  DevBuf<uint32_t> perm;
  DevBuf<uint32_t> status;  // [0] duplicate flag, [1] Morton fallback count, [2..] scratch

  // global AABB: lo[3], hi[3] values and PointId witnesses
  DevBuf<double> bounds;
  DevBuf<uint32_t> bound_ids;

  // LBVH: leaves in (Morton code, PointId) order
  DevBuf<uint64_t> codes;
  DevBuf<uint32_t> leaf_id;   // Morton position -> PointId
  DevBuf<uint32_t> leaf_pos;  // PointId -> Morton position
  DevBuf<double> LX, LY, LZ;  // coordinates in Morton order
  DevBuf<uint64_t> path;      // midpoint path bits inside equal-code runs
  DevBuf<Node> nodes;         // n-1 internal nodes, root = 0
  DevBuf<uint32_t> parent_int, parent_leaf;
  DevBuf<uint32_t> wit;       // 6 witnesses per internal node (lo xyz, hi xyz)
  DevBuf<WideNode> wide;      // FP32 two-level copy of the internal nodes (ball and k-NN queries)
  DevBuf<uint32_t> counter;
  RadixSorter sorter;
  Scanner scanner;
  mhgp_stats stats{};

  // device time of one launch window on the context stream
  struct KernelClock {
    hipEvent_t a = nullptr, b = nullptr;
    void start(hipStream_t s) {
      if (!a) { MHGP_HIP(hipEventCreate(&a)); MHGP_HIP(hipEventCreate(&b)); }
      MHGP_HIP(hipEventRecord(a, s));
    }
    void stop(hipStream_t s) { MHGP_HIP(hipEventRecord(b, s)); }
    double ms() {
      float v = 0;
      MHGP_HIP(hipEventSynchronize(b));
      MHGP_HIP(hipEventElapsedTime(&v, a, b));
      return v;
    }
  } anchor_clock, arm_clock;

  std::unique_ptr<CatalogBuffers> catalog;
  std::unique_ptr<EmstBuffers> emst;
  std::unique_ptr<HostMirror> host;
  std::unique_ptr<TowerState> tower;
  std::unique_ptr<TowerWork> tower_work;                 // first lane of the tower (context stream)
  std::vector<std::unique_ptr<TowerWork>> tower_lanes;  // further lanes (own streams)
  PinnedBuf tower_pinned[11];  // per tower order: saddle results on the host

  Context();
  ~Context();
};

// lbvh.hip
void canonicalize_and_build(Context &c);
void export_snapshot(Context &c, uint64_t *leaves, uint64_t *nodes, uint64_t *counters);

}  // namespace mhgp
