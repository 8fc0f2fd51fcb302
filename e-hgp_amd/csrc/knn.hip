// k nearest neighbours (k <= 32) of every data point by LBVH traversal.
//
// Semantics follow the reference's exact top-k (morsehgp3d/src/cpu/spatial/
// lbvh.cpp:1098-1198 lbvh_top_k): neighbours ordered by exact squared distance,
// ties ordered by PointId; the query point itself is excluded.  One thread
// per query, queries taken in Morton order so that neighbouring lanes walk
// the same subtrees.  Distances are FP64; a pair of candidates whose values
// are within rounding of each other and not both exact is reported in
// tie_flags so that the host can settle the order exactly.
#include "context.hpp"
#include "traverse.hpp"

namespace mhgp {

namespace {

constexpr int KMAX = 33;  // k + 1 slots for k <= 32

__global__ void __launch_bounds__(128) k_knn(int64_t n, int k, const Node *__restrict__ nodes,
                                             const double *__restrict__ LX,
                                             const double *__restrict__ LY,
                                             const double *__restrict__ LZ,
                                             const uint32_t *__restrict__ leaf_id,
                                             uint32_t *out, uint8_t *flags) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int kk = k + 1 < (int)n ? k + 1 : (int)n - 1;  // slots kept (k, plus one tie probe)
  const uint32_t self = leaf_id[p];
  const double q[3] = {LX[p], LY[p], LZ[p]};
  double bd[KMAX];
  uint32_t bi[KMAX];
  bool bx[KMAX];
  int cnt = 0;
  double cut = INFINITY;  // squared distance of the kk-th best so far
  uint32_t stack[64];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    uint32_t nd = stack[--sp];
    const Node &N = nodes[nd];
    uint32_t ch[2] = {N.left, N.right};
    double cd[2];
    for (int c = 0; c < 2; ++c) {
      uint32_t x = ch[c];
      if (x & kLeafBit) {
        uint32_t lp = x & ~kLeafBit;
        cd[c] = INFINITY;
        uint32_t id = leaf_id[lp];
        if (id == self) continue;
        bool ex;
        double d = d2_flag(q, LX[lp], LY[lp], LZ[lp], ex);
        if (cnt == kk && (d > cut || (d == cut && id > bi[kk - 1]))) continue;
        // insertion by (d, id)
        int pos = cnt < kk ? cnt : kk - 1;
        while (pos > 0 && (bd[pos - 1] > d || (bd[pos - 1] == d && bi[pos - 1] > id))) {
          bd[pos] = bd[pos - 1]; bi[pos] = bi[pos - 1]; bx[pos] = bx[pos - 1];
          pos--;
        }
        bd[pos] = d; bi[pos] = id; bx[pos] = ex;
        if (cnt < kk) cnt++;
        if (cnt == kk) cut = bd[kk - 1];
      } else {
        const Node &C = nodes[x];
        cd[c] = box_d2(q, C.lo, C.hi);
      }
    }
    // push farther child first so the nearer one is visited next
    int first = cd[0] <= cd[1] ? 0 : 1;
    for (int o = 1; o >= 0; --o) {
      int c = o ? 1 - first : first;
      uint32_t x = ch[c];
      if (x & kLeafBit) continue;
      if (cd[c] > cut * (1.0 + kSlack) + kSlackAbs) continue;
      if (sp < 64) stack[sp++] = x;
      else flags[self] |= 4;  // stack overflow: host recomputes
    }
  }
  uint8_t f = 0;
  for (int i = 0; i + 1 < cnt; ++i) {
    double a = bd[i], b = bd[i + 1];
    if (!(bx[i] && bx[i + 1]) && (b - a) <= kSlack * b + kSlackAbs) f |= 1;
  }
  const int64_t o = (int64_t)self * k;
  for (int i = 0; i < k; ++i) out[o + i] = i < cnt ? bi[i] : kInvalid;
  if (f) flags[self] |= f;
}

}  // namespace

void knn_all(Context &c, int k, uint32_t *d_out, uint8_t *d_flags) {
  if (k < 1 || k > 32) throw std::invalid_argument("k must lie in 1..32");
  if (k >= c.n) throw std::invalid_argument("k must be smaller than the point count");
  MHGP_HIP(hipMemsetAsync(d_flags, 0, c.n, c.stream));
  hipLaunchKernelGGL(k_knn, dim3(blocks_for(c.n, 128)), dim3(128), 0, c.stream, c.n, k, c.nodes.p,
                     c.LX.p, c.LY.p, c.LZ.p, c.leaf_id.p, d_out, d_flags);
  MHGP_LAUNCH_CHECK();
}

}  // namespace mhgp
