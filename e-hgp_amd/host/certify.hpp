// Exact re-certification of an LBVH snapshot (host only).
#pragma once

#include <cstdint>
#include <string>

namespace mhgp {

// Empty when the snapshot (reference layout: leaves 2n u64, nodes 10 u64 x
// (2n - 1), counters 5 u64) is exactly the Morton LBVH of the canonical cloud
// xyz (n x 3 binary64, PointId order); otherwise the first violation found.
std::string certify_snapshot(const double *xyz, int64_t n, const uint64_t *leaves, const uint64_t *nodes,
                             const uint64_t *counters);

}  // namespace mhgp
