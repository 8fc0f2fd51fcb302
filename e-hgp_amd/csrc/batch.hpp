// Batched catalogue: one persistent anchor launch over several clouds.
#pragma once

#include <cstdint>
#include <memory>

namespace mhgp {

struct AnchorBatch;  // supports.hip

struct AnchorBatchStats {
  double ms = 0;        // device time of the persistent launch
  int64_t anchors = 0;  // anchors it processed
  int64_t jobs = 0;     // clouds in it
};

// A rendezvous for `participants` clouds on one device.  Each cloud's
// context points at it (Context::batch, Context::batch_slot) while its
// catalogue runs; a cloud that will not reach its first anchor pass leaves.
std::shared_ptr<AnchorBatch> anchor_batch_create(int device, int participants);
void anchor_batch_leave(AnchorBatch &b, int slot);
AnchorBatchStats anchor_batch_stats(AnchorBatch &b);

}  // namespace mhgp
