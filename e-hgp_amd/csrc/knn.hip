// k nearest neighbours (k <= 32) of every data point by LBVH traversal.
//
// Semantics follow the reference's exact top-k (morsehgp3d/src/cpu/spatial/
// lbvh.cpp:1098-1198 lbvh_top_k): neighbours ordered by exact squared distance,
// ties ordered by PointId; the query point itself is excluded.  One thread
// per query, queries taken in Morton order so that neighbouring lanes walk
// the same subtrees.  Distances are FP64; a pair of candidates whose values
// are within rounding of each other and not both exact is reported in
// tie_flags so that the host can settle the order exactly.
#include <algorithm>
#include <cmath>

#include "../host/exact.hpp"
#include "context.hpp"
#include "pipeline.hpp"
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

// Exact top-k partition of rational query points (reference lbvh_top_k,
// morsehgp3d/include/morsehgp3d/spatial/lbvh.hpp:424 with the TopKPartition
// of spatial/query.hpp:109-169).  The device proposes: per query, the FP64
// k-th smallest distance r_k over the eligible points from the rounded query
// (pass 1), then every eligible point within r_k widened by the query's
// rounding and the distance roundings (pass 2).  Every point whose exact
// distance is at most the exact k-th lies in that set, so the host's exact
// replay over it decides the cutoff, the strict-below list and the complete
// cutoff shell.  Queries whose set overflows, or whose exclusion list or
// traversal stack does not fit, are settled by an exact host scan.
constexpr int kPartCand = 64;   // candidate slots per query
constexpr int kPartExcl = 64;   // exclusions checked on the device per query
constexpr uint32_t kPartHost = 1u;

struct PartParams {
  const double *q;        // 3 per query, device units
  const double *slack;    // per query: absolute distance slack (device units)
  const uint32_t *excl;   // exclusion ids, CSR
  const int64_t *eoff;
  int64_t queries;
  int k;
  int64_t n;
  const Node *nodes;
  const double *LX, *LY, *LZ;
  const uint32_t *leaf_id;
  uint32_t *cand;         // kPartCand per query
  uint32_t *ncand;
  uint32_t *flags;
};

__device__ inline bool part_excluded(const PartParams &P, int64_t q, uint32_t id) {
  for (int64_t j = P.eoff[q]; j < P.eoff[q + 1]; ++j)
    if (P.excl[j] == id) return true;
  return false;
}

__global__ void __launch_bounds__(128) k_knn_partition(PartParams P) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P.queries) return;
  P.ncand[q] = 0;
  P.flags[q] = 0;
  if (P.eoff[q + 1] - P.eoff[q] > kPartExcl) {
    P.flags[q] = kPartHost;
    return;
  }
  const double qq[3] = {P.q[3 * q], P.q[3 * q + 1], P.q[3 * q + 2]};
  const int k = P.k;
  double bd[32];
  int cnt = 0;
  double cut = INFINITY;
  uint32_t stack[64];
  int sp = 0;
  bool overflow = false;
  stack[sp++] = 0;
  // pass 1: the FP64 k-th smallest squared distance over eligible points
  while (sp) {
    const Node &N = P.nodes[stack[--sp]];
    const uint32_t ch[2] = {N.left, N.right};
    for (int c = 0; c < 2; ++c) {
      const uint32_t x = ch[c];
      if (x & kLeafBit) {
        const uint32_t lp = x & ~kLeafBit;
        const double dx = P.LX[lp] - qq[0], dy = P.LY[lp] - qq[1], dz = P.LZ[lp] - qq[2];
        const double d = dx * dx + dy * dy + dz * dz;
        if (cnt == k && !(d < cut)) continue;
        if (part_excluded(P, q, P.leaf_id[lp])) continue;
        int pos = cnt < k ? cnt : k - 1;
        while (pos > 0 && bd[pos - 1] > d) { bd[pos] = bd[pos - 1]; pos--; }
        bd[pos] = d;
        if (cnt < k) cnt++;
        if (cnt == k) cut = bd[k - 1];
      } else if (box_d2(qq, P.nodes[x].lo, P.nodes[x].hi) <= cut * (1.0 + kSlack) + kSlackAbs) {
        if (sp < 64) stack[sp++] = x;
        else overflow = true;
      }
    }
  }
  if (overflow || cnt < k) {
    P.flags[q] = kPartHost;
    return;
  }
  // pass 2: every eligible point that can be at or below the exact k-th
  const double rk = sqrt(cut);
  const double R = (rk * (1.0 + 1e-12) + 2.0 * P.slack[q]) * (1.0 + 1e-12) + 1e-300;
  const double R2 = R * R * (1.0 + 1e-12);
  uint32_t nc = 0;
  sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const Node &N = P.nodes[stack[--sp]];
    const uint32_t ch[2] = {N.left, N.right};
    for (int c = 0; c < 2; ++c) {
      const uint32_t x = ch[c];
      if (x & kLeafBit) {
        const uint32_t lp = x & ~kLeafBit;
        const double dx = P.LX[lp] - qq[0], dy = P.LY[lp] - qq[1], dz = P.LZ[lp] - qq[2];
        if (dx * dx + dy * dy + dz * dz > R2) continue;
        const uint32_t id = P.leaf_id[lp];
        if (part_excluded(P, q, id)) continue;
        if (nc < kPartCand) P.cand[q * kPartCand + nc] = id;
        nc++;
      } else if (box_d2(qq, P.nodes[x].lo, P.nodes[x].hi) <= R2 * (1.0 + kSlack) + kSlackAbs) {
        if (sp < 64) stack[sp++] = x;
        else overflow = true;
      }
    }
  }
  P.ncand[q] = nc;
  if (overflow || nc > (uint32_t)kPartCand) P.flags[q] = kPartHost;
}

}  // namespace

void knn_partition(Context &c, const int64_t *num, const int64_t *den, int64_t Q, int k, const uint32_t *excl,
                   const int64_t *eoff, uint32_t shell_cap, uint32_t *choice, uint32_t *strict,
                   uint32_t *strict_count, uint32_t *shell, uint32_t *shell_count, uint32_t *cutoff_id,
                   double *cutoff_sq) {
  require_cloud(c);
  const int64_t n = c.n;
  if (k < 1 || k > 32) throw std::invalid_argument("requested rank must lie in 1..32");
  if (Q < 0 || (Q > 0 && (!num || !den || !choice || !strict || !strict_count || !shell_count || !cutoff_id ||
                          !cutoff_sq || (shell_cap && !shell))))
    throw std::invalid_argument("invalid top-k partition arrays");
  if (Q == 0) return;
  std::vector<int64_t> off(Q + 1, 0);
  if (eoff) {
    if (eoff[0] != 0) throw std::invalid_argument("exclusion offsets must start at zero");
    for (int64_t q = 0; q < Q; ++q) {
      if (eoff[q + 1] < eoff[q]) throw std::invalid_argument("exclusion offsets must not decrease");
      off[q + 1] = eoff[q + 1];
    }
  }
  const int64_t E = off[Q];
  if (E && !excl) throw std::invalid_argument("null exclusion ids");
  std::vector<double> qf(3 * (size_t)Q), slack(Q);
  for (int64_t q = 0; q < Q; ++q) {
    if (den[q] <= 0) throw std::invalid_argument("query denominators must be positive");
    // eligible points: distinct, in range, not excluded
    std::vector<uint32_t> ex(excl + off[q], excl + off[q + 1]);
    std::sort(ex.begin(), ex.end());
    ex.erase(std::unique(ex.begin(), ex.end()), ex.end());
    for (uint32_t id : ex)
      if (id >= (uint64_t)n) throw std::invalid_argument("excluded PointId out of range");
    if ((int64_t)ex.size() + k > n) throw std::invalid_argument("the requested rank exceeds the eligible point count");
    double m = 0;
    for (int a = 0; a < 3; ++a) {
      // num / den rounded (each conversion and the division lose at most half
      // an ulp), in the device's unit
      const double v = (double)num[3 * q + a] / (double)den[q];
      qf[3 * q + a] = std::ldexp(v, c.scale_exp);
      m = std::max(m, std::fabs(v));
    }
    slack[q] = std::ldexp(m * 8.0 * 0x1p-53 * 1.7320508075688772 + 1e-300, c.scale_exp) + 1e-300;
  }
  hipStream_t s = c.stream;
  DevBuf<double> d_q, d_slack;
  DevBuf<uint32_t> d_ex, d_cand, d_nc, d_fl;
  DevBuf<int64_t> d_off;
  d_q.ensure(3 * Q);
  d_slack.ensure(Q);
  d_ex.ensure(std::max<int64_t>(E, 1));
  d_off.ensure(Q + 1);
  d_cand.ensure((size_t)Q * kPartCand);
  d_nc.ensure(Q);
  d_fl.ensure(Q);
  MHGP_HIP(hipMemcpyAsync(d_q.p, qf.data(), 3 * Q * 8, hipMemcpyHostToDevice, s));
  MHGP_HIP(hipMemcpyAsync(d_slack.p, slack.data(), Q * 8, hipMemcpyHostToDevice, s));
  if (E) MHGP_HIP(hipMemcpyAsync(d_ex.p, excl, E * 4, hipMemcpyHostToDevice, s));
  MHGP_HIP(hipMemcpyAsync(d_off.p, off.data(), (Q + 1) * 8, hipMemcpyHostToDevice, s));
  std::vector<uint32_t> cand((size_t)Q * kPartCand), nc(Q), fl(Q, kPartHost);
  if (n > 1) {
    PartParams P{d_q.p, d_slack.p, d_ex.p, d_off.p, Q, k, n, c.nodes.p, c.LX.p, c.LY.p, c.LZ.p, c.leaf_id.p,
                 d_cand.p, d_nc.p, d_fl.p};
    hipLaunchKernelGGL(k_knn_partition, dim3(blocks_for(Q, 128)), dim3(128), 0, s, P);
    MHGP_LAUNCH_CHECK();
    MHGP_HIP(hipMemcpyAsync(cand.data(), d_cand.p, cand.size() * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(nc.data(), d_nc.p, Q * 4, hipMemcpyDeviceToHost, s));
    copy_wait(fl.data(), d_fl.p, Q * 4, hipMemcpyDeviceToHost, s);
  }
  // exact replay: squared distances times den^2 (and a power of two common to
  // the whole cloud) as integers, X_a = P_a * den - num_a scaled to integers
  const ExactCloud &ex = exact_cloud(c);
  const int e0 = ex.e0();
  std::vector<std::pair<BigInt, uint32_t>> ev;
  for (int64_t q = 0; q < Q; ++q) {
    BigInt bden(den[q]), bnum[3];
    for (int a = 0; a < 3; ++a) bnum[a] = BigInt(num[3 * q + a]);
    // value = P 2^e0: with e0 < 0 the numerators take the power instead
    BigInt pden = e0 >= 0 ? BigInt::from_dyadic(den[q], e0) : bden;
    if (e0 < 0)
      for (int a = 0; a < 3; ++a) bnum[a] = bnum[a] * BigInt::from_dyadic(1, -e0);
    std::vector<uint32_t> exq(excl + off[q], excl + off[q + 1]);
    std::sort(exq.begin(), exq.end());
    auto excluded = [&](uint32_t id) { return std::binary_search(exq.begin(), exq.end(), id); };
    auto dist = [&](uint32_t id) {
      BigInt sum(0);
      for (int a = 0; a < 3; ++a) {
        BigInt x = ex.coord(id, a) * pden - bnum[a];
        sum = sum + x * x;
      }
      return sum;
    };
    ev.clear();
    if (fl[q] & kPartHost) {
      for (int64_t id = 0; id < n; ++id)
        if (!excluded((uint32_t)id)) ev.emplace_back(dist((uint32_t)id), (uint32_t)id);
    } else {
      for (uint32_t i = 0; i < nc[q]; ++i) {
        const uint32_t id = cand[(size_t)q * kPartCand + i];
        ev.emplace_back(dist(id), id);
      }
    }
    if ((int64_t)ev.size() < k) throw std::logic_error("top-k partition: fewer candidates than the requested rank");
    std::sort(ev.begin(), ev.end(), [](const auto &a, const auto &b) {
      const int s2 = compare(a.first, b.first);
      return s2 != 0 ? s2 < 0 : a.second < b.second;
    });
    const BigInt &cutv = ev[k - 1].first;
    uint32_t ns = 0, nsh = 0;
    std::vector<uint32_t> sh;
    for (const auto &e : ev) {
      const int s2 = compare(e.first, cutv);
      if (s2 < 0) strict[(size_t)q * k + ns++] = e.second;
      else if (s2 == 0) sh.push_back(e.second);
    }
    std::sort(sh.begin(), sh.end());
    nsh = (uint32_t)sh.size();
    for (uint32_t i = ns; i < (uint32_t)k; ++i) strict[(size_t)q * k + i] = kInvalid;
    strict_count[q] = ns;
    shell_count[q] = nsh;
    for (uint32_t i = 0; i < shell_cap; ++i) shell[(size_t)q * shell_cap + i] = i < nsh ? sh[i] : kInvalid;
    cutoff_id[q] = sh[0];
    std::vector<uint32_t> ch(strict + (size_t)q * k, strict + (size_t)q * k + ns);
    for (uint32_t i = 0; ns + i < (uint32_t)k; ++i) ch.push_back(sh[i]);
    std::sort(ch.begin(), ch.end());
    std::copy(ch.begin(), ch.end(), choice + (size_t)q * k);
    // cutoff = sum / den^2 in input units (times 2^(2 e0) when e0 < 0)
    cutoff_sq[q] = cutv.sign() == 0 ? 0.0 : ratio_to_double(cutv, bden * bden, e0 < 0 ? 2 * e0 : 0);
  }
}

void knn_all(Context &c, int k, uint32_t *d_out, uint8_t *d_flags, hipStream_t stream) {
  if (k < 1 || k > 32) throw std::invalid_argument("k must lie in 1..32");
  if (k >= c.n) throw std::invalid_argument("k must be smaller than the point count");
  hipStream_t s = stream ? stream : c.stream;
  MHGP_HIP(hipMemsetAsync(d_flags, 0, c.n, s));
  hipLaunchKernelGGL(k_knn, dim3(blocks_for(c.n, 128)), dim3(128), 0, s, c.n, k, c.nodes.p, c.LX.p, c.LY.p, c.LZ.p,
                     c.leaf_id.p, d_out, d_flags);
  MHGP_LAUNCH_CHECK();
}

}  // namespace mhgp
