// Reference-side binding check, compiled on the host with g++ against
// include/mhgp.h and linked with libmhgp.so.
//
// It mirrors the reference's snapshot records (MortonLbvhSnapshotLeaf, 16 B,
// and MortonLbvhSnapshotNode, 80 B; morsehgp3d/include/morsehgp3d/spatial/
// lbvh.hpp:43-103) with static_assert-ed layouts, fills them directly from
// mhgp_lbvh_snapshot (no repacking), and checks what the reference's
// certified import checks structurally (lbvh.cpp:562-829): 2n-1 nodes in
// strict postorder with the root last, children before parents, contiguous
// leaf ranges that partition [0, n), leaves with no children, and Morton
// codes non-decreasing in leaf order.  Prints "ok" and the counters.
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mhgp.h"

namespace mirror {

using PointId = std::uint64_t;
constexpr std::uint64_t kInvalidNode = ~std::uint64_t{0};

struct MortonLbvhSnapshotLeaf {
  std::uint64_t morton_code;
  PointId point_id;
};
static_assert(sizeof(MortonLbvhSnapshotLeaf) == 16, "leaf record is 16 bytes");
static_assert(offsetof(MortonLbvhSnapshotLeaf, point_id) == 8, "leaf layout");

struct MortonLbvhSnapshotNode {
  PointId lower_point_ids[3];
  PointId upper_point_ids[3];
  std::uint64_t left_child;
  std::uint64_t right_child;
  std::uint64_t leaf_begin;
  std::uint64_t leaf_end;
};
static_assert(sizeof(MortonLbvhSnapshotNode) == 80, "node record is 80 bytes");
static_assert(offsetof(MortonLbvhSnapshotNode, left_child) == 48, "node layout");
static_assert(offsetof(MortonLbvhSnapshotNode, leaf_end) == 72, "node layout");

}  // namespace mirror

static int fail(const char *what) {
  std::printf("FAIL %s: %s\n", what, mhgp_last_error());
  return 1;
}

int main(int argc, char **argv) {
  const int64_t n = argc > 1 ? std::atoll(argv[1]) : 4096;
  std::vector<double> xyz(3 * n);
  std::uint64_t s = 0x9E3779B97F4A7C15ull;
  for (auto &v : xyz) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    v = (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
  mhgp_ctx *ctx = nullptr;
  if (mhgp_ctx_create(0, 0, &ctx) != MHGP_OK) return fail("ctx_create");
  if (mhgp_cloud_upload(ctx, xyz.data(), n) != MHGP_OK) return fail("upload");
  std::vector<mirror::MortonLbvhSnapshotLeaf> leaves(n);
  std::vector<mirror::MortonLbvhSnapshotNode> nodes(2 * n - 1);
  std::uint64_t counters[5];
  if (mhgp_lbvh_snapshot(ctx, reinterpret_cast<std::uint64_t *>(leaves.data()),
                         reinterpret_cast<std::uint64_t *>(nodes.data()), counters) != MHGP_OK)
    return fail("snapshot");
  mhgp_ctx_destroy(ctx);
  if (counters[0] != (std::uint64_t)n || counters[1] != (std::uint64_t)(2 * n - 1)) return fail("counters");
  std::vector<char> seen(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (leaves[i].point_id >= (std::uint64_t)n || seen[leaves[i].point_id]) return fail("leaf permutation");
    seen[leaves[i].point_id] = 1;
    if (i && leaves[i].morton_code < leaves[i - 1].morton_code) return fail("leaf order");
  }
  for (std::size_t i = 0; i < nodes.size(); ++i) {
    const auto &nd = nodes[i];
    if (nd.left_child == mirror::kInvalidNode) {
      if (nd.right_child != mirror::kInvalidNode || nd.leaf_end != nd.leaf_begin + 1) return fail("leaf node");
      continue;
    }
    if (nd.left_child >= i || nd.right_child >= i) return fail("postorder");
    const auto &l = nodes[nd.left_child], &r = nodes[nd.right_child];
    if (l.leaf_begin != nd.leaf_begin || l.leaf_end != r.leaf_begin || r.leaf_end != nd.leaf_end)
      return fail("leaf ranges");
  }
  const auto &root = nodes.back();
  if (root.leaf_begin != 0 || root.leaf_end != (std::uint64_t)n) return fail("root range");
  std::printf("ok n=%lld depth=%llu collision_groups=%llu\n", (long long)n, (unsigned long long)counters[2],
              (unsigned long long)counters[3]);
  return 0;
}
