#pragma once

#include <cstdint>
#include <initializer_list>
#include <vector>

#include "context.hpp"
#include "radix_sort.hpp"
#include "scan.hpp"

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
  std::vector<uint32_t> raw_of_dense;   // raw minimum of each dense contraction element (empty: identity)
  uint32_t vm_first_pos = 0;            // rank-index position of the first saddle
  int64_t steps = 0;
};

// Device scratch of the per-order tower work, kept by the context; a buffer
// is reallocated only when an order needs more than it holds.
struct TowerWork {
  // the lane's stream (the context's for the first lane), scan and sort
  // scratch, kernel clock and statistics
  hipStream_t stream = nullptr;
  bool own_stream = false;
  Scanner scanner;
  RadixSorter sorter;
  Context::KernelClock clock;
  double arm_ms = 0;
  int64_t arm_launches = 0, arm_bytes = 0;
  TowerWork() = default;
  TowerWork(const TowerWork &) = delete;
  TowerWork &operator=(const TowerWork &) = delete;
  ~TowerWork() {
    if (vm_stream) (void)hipStreamSynchronize(vm_stream);
    if (nn_start) (void)hipEventDestroy(nn_start);
    if (nn_done) (void)hipEventDestroy(nn_done);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    if (vm_stream) (void)hipStreamDestroy(vm_stream);
  }
  // the side stream of the first lane: the neighbour lists, computed beside
  // the catalogue's level order, and the vertical maps' element gathers
  hipStream_t vm_stream = nullptr;
  hipEvent_t nn_start = nullptr, nn_done = nullptr;
  bool nn_pending = false;  // lists launched on the side stream for the current cloud
  // records grouped by closed rank: rank r occupies rank_vals[rank_lo[r], rank_hi[r])
  DevBuf<uint64_t> rank_keys;
  DevBuf<uint32_t> rank_vals, rank_bounds;
  uint32_t rank_lo[256] = {}, rank_hi[256] = {};
  DevBuf<uint32_t> tvals, labels;
  DevBuf<uint64_t> tkeys, steps, qcount, qkey, facet_slots, alias_count;
  DevBuf<uint32_t> alias_arm, alias_rep, dF[2], dH[2], dO[2];
  DevBuf<uint64_t> dcount;
  DevBuf<uint32_t> sizes, carrier, scls, base, arm_saddle, queue;
  DevBuf<double> slev, c_lev;
  DevBuf<uint32_t> first, keep, kept, flag1, c_cls, c_base, c_car;
  DevBuf<uint32_t> eflag, eu, ev, es;
  DevBuf<uint32_t> comp, comp2, best, hook, msf, mcount, changed, err, live_n;
  DevBuf<uint32_t> Lu[2], Lv[2], Li[2];
  DevBuf<uint32_t> attach, rep, dense, slab, ssup;
  // vertical maps: per record (rank-index position) the contraction element
  // of its first arm's carrier as a saddle of the order below (first lane
  // only), the gather scratch, and the dense -> raw minimum map of a lane
  DevBuf<uint32_t> vm_elem, vm_idx, vm_out, raw_dense;
  // nearest-neighbour lists of every point and the squared distance to the
  // last listed one (first lane only)
  DevBuf<uint32_t> nn;
  DevBuf<double> nn_d2;
  DevBuf<uint8_t> nn_flags;
  bool nn_ready = false;
  int64_t bytes() const {
    int64_t b = 0;
    for (const DevBuf<uint32_t> *x : {&rank_vals, &rank_bounds, &tvals, &labels, &sizes, &carrier, &scls,
                                      &base, &arm_saddle, &queue, &first, &keep, &kept, &flag1, &c_cls, &c_base,
                                      &c_car, &eflag, &eu, &ev, &es, &comp, &comp2, &best, &hook, &msf, &mcount,
                                      &changed, &err, &live_n, &Lu[0], &Lu[1], &Lv[0], &Lv[1], &Li[0], &Li[1],
                                      &attach, &rep, &dense, &slab, &ssup, &vm_elem, &vm_idx, &vm_out, &raw_dense, &alias_arm, &alias_rep, &dF[0], &dF[1], &dH[0], &dH[1], &dO[0], &dO[1]})
      b += (int64_t)x->bytes();
    return b + (int64_t)(nn.bytes() + nn_d2.bytes() + nn_flags.bytes()) + (int64_t)(rank_keys.bytes() + qkey.bytes() + facet_slots.bytes() + alias_count.bytes() + dcount.bytes() + tkeys.bytes() + steps.bytes() + qcount.bytes() + slev.bytes() + c_lev.bytes());
  }
};

// Starts the neighbour lists of the descents on the side stream (after the
// LBVH, before the catalogue's level order), so that they overlap it.
void tower_prefetch_neighbours(Context &c);
// Groups the catalogue's records by closed rank (once per tower call).
void tower_rank_index(Context &c);
// Resolves the arms of order k and reduces its saddles (after
// tower_rank_index) on the lane W; R holds the rank index.
void tower_order_device(Context &c, TowerWork &W, const TowerWork &R, int k, TowerOrderData &out);
// Vertical maps: contraction elements at the given rank-index positions
// (the first lane's gather stream and scratch; one caller at a time), and
// host-settled elements written back.
void tower_vm_gather(Context &c, const std::vector<uint32_t> &pos, std::vector<uint32_t> &out);
void tower_vm_store(Context &c, hipStream_t s, uint32_t first_pos, const std::vector<uint32_t> &elems);

}  // namespace mhgp
