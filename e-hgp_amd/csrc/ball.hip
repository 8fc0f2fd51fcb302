// Closed-ball partition of the cloud against the circumsphere of a support
// (reference lbvh_closed_ball, morsehgp3d/src/cpu/spatial/lbvh.cpp:1210-1340,
// and the sparse classification of the support streams,
// pair_support_stream.cpp:1781-1952): how many points lie strictly inside
// the sphere, how many on it (the support included) and which are inside.
//
// One thread per query: the FP64 circumball bounds an LBVH ball traversal
// (with a relative margin far above its rounding error), every point in range
// is placed by the certified predicates of predicates.hpp, and what they
// cannot decide (points exactly or nearly on the sphere, ill-conditioned
// supports) is settled on the host in exact arithmetic.
#include <algorithm>
#include <mutex>

#include "../host/exact.hpp"
#include "pipeline.hpp"
#include "predicates.hpp"
#include "traverse.hpp"

namespace mhgp {

namespace {

constexpr int kBallStack = 40;  // LBVH depth is 18 at 50k points, 32 at 200k; overflow goes to the host
constexpr uint32_t kHostQuery = 1u;

struct BallParams {
  const uint32_t *ids;    // 4 per query
  const uint8_t *sizes;   // support size per query (1..4)
  int64_t queries;
  uint32_t cap;           // interior ids kept per query
  const double *X, *Y, *Z;
  const Node *nodes;
  const double *LX, *LY, *LZ;
  const uint32_t *leaf_id;
  int64_t n;
  uint32_t *interior;     // cap per query
  uint32_t *nin, *nshell, *flags;
  uint32_t *und;          // (query, point) pairs for the host
  unsigned long long *nund;
  uint64_t und_cap;
};

__device__ inline V3 pt(const BallParams &B, uint32_t id) { return {B.X[id], B.Y[id], B.Z[id]}; }

__global__ void __launch_bounds__(64) k_closed_ball(BallParams B) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B.queries) return;
  const int m = B.sizes[q];
  uint32_t sid[4];
  for (int i = 0; i < 4; ++i) sid[i] = i < m ? B.ids[4 * q + i] : kInvalid;
  B.nin[q] = 0;
  B.nshell[q] = (uint32_t)m;
  B.flags[q] = 0;
  if (m == 1) {  // radius zero: nothing inside, the point alone on it
    return;
  }
  // circumsphere: centre o + off in the support's affine hull
  const V3 o = pt(B, sid[0]);
  const V3 a = sub(pt(B, sid[1]), o);
  PairSphere ps;
  TriSphere ts;
  TetSphere tt;
  V3 off;
  double r2;
  bool good;
  if (m == 2) {
    ps.init(a);
    off = {0.5 * a.x, 0.5 * a.y, 0.5 * a.z};
    good = dot(a, a) > 0;
  } else if (m == 3) {
    const V3 b = sub(pt(B, sid[2]), o);
    ts.init(a, b);
    good = ts.D > 1e-6 * ts.DA;
    const double s2 = good ? 0.5 / ts.D : 0.0;
    off = {ts.g.x * s2, ts.g.y * s2, ts.g.z * s2};
  } else {
    const V3 b = sub(pt(B, sid[2]), o), d = sub(pt(B, sid[3]), o);
    tt.init(a, b, d);
    good = fabs(tt.V) > 1e-6 * tt.VA;
    const double s2 = good ? 0.5 / tt.V : 0.0;
    off = {tt.N.x * s2, tt.N.y * s2, tt.N.z * s2};
  }
  if (!good) {  // dependent or ill-conditioned support: the host classifies it
    B.flags[q] = kHostQuery;
    return;
  }
  r2 = dot(off, off);
  const double c[3] = {o.x + off.x, o.y + off.y, o.z + off.z};
  const double R2 = r2 * (1.0 + 1e-6) + 1e-300;  // every point beyond is certainly outside
  uint32_t nin = 0, nsh = (uint32_t)m;
  // the interior row keeps the cap smallest ids (their largest and its slot
  // tracked), so a truncated list does not depend on the traversal order
  uint32_t *row = B.interior + (size_t)q * B.cap;
  uint32_t rmax = 0, rmax_at = 0;
  uint32_t stack[kBallStack];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const Node &N = B.nodes[stack[--sp]];
    const uint32_t ch[2] = {N.left, N.right};
    for (int cc = 0; cc < 2; ++cc) {
      const uint32_t x = ch[cc];
      if (x & kLeafBit) {
        const uint32_t lp = x & ~kLeafBit;
        const double dx = B.LX[lp] - c[0], dy = B.LY[lp] - c[1], dz = B.LZ[lp] - c[2];
        if (dx * dx + dy * dy + dz * dz > R2) continue;
        const uint32_t id = B.leaf_id[lp];
        bool on_support = false;
        for (int i = 0; i < m; ++i) on_support |= sid[i] == id;
        if (on_support) continue;
        const V3 p = sub(pt(B, id), o);
        const int sd = m == 2 ? ps.side(p) : (m == 3 ? ts.side(p) : tt.side(p));
        if (sd < 0) {
          if (nin < B.cap) {
            row[nin] = id;
            if (id >= rmax) { rmax = id; rmax_at = nin; }
          } else if (B.cap && id < rmax) {
            row[rmax_at] = id;
            rmax = 0;
            for (uint32_t k = 0; k < B.cap; ++k)
              if (row[k] >= rmax) { rmax = row[k]; rmax_at = k; }
          }
          nin++;
        } else if (sd == 0) {
          const unsigned long long at = atomicAdd(B.nund, 1ull);
          if (at < B.und_cap) {
            B.und[2 * at] = (uint32_t)q;
            B.und[2 * at + 1] = id;
          } else {
            B.flags[q] = kHostQuery;  // list full: the whole query goes to the host
          }
        }
      } else if (box_d2(c, B.nodes[x].lo, B.nodes[x].hi) <= R2) {
        if (sp < kBallStack) stack[sp++] = x;
        else B.flags[q] = kHostQuery;
      }
    }
  }
  B.nin[q] = nin;
  B.nshell[q] = nsh;
}

// Allocation-free variant: the host-query flag of each query is folded into
// the high bit of its interior count (counts stay below 2^31 since n does),
// so the host reads both with one copy into the caller's own array.
__global__ void __launch_bounds__(256) k_fold_flags(uint32_t *nin, const uint32_t *flags, int64_t queries) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < queries && (flags[q] & kHostQuery)) nin[q] |= 0x80000000u;
}

__global__ void k_zero_u64(unsigned long long *x) { *x = 0ull; }

// The undecided (query, point) pairs as sortable keys query << 32 | point,
// padded with the largest key up to the power of two the sort works on.
__global__ void k_und_keys(uint64_t *und, uint64_t count, uint64_t padded) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= padded) return;
  if (i < count) {
    const uint32_t *w = reinterpret_cast<const uint32_t *>(und + i);
    und[i] = ((uint64_t)w[0] << 32) | w[1];
  } else {
    und[i] = ~0ull;
  }
}

// One compare-exchange stage (size k, distance j) of a bitonic sort, ascending.
__global__ void k_bitonic_stage(uint64_t *a, uint64_t n, uint64_t j, uint64_t k) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t l = i ^ j;
  if (i >= n || l <= i) return;
  const uint64_t x = a[i], y = a[l];
  const bool up = (i & k) == 0;
  if (up ? x > y : x < y) {
    a[i] = y;
    a[l] = x;
  }
}

// Device scratch of mhgp_closed_ball_noalloc: query staging, per-query
// counts and flags, and the list of undecided (query, point) pairs.  Each
// device holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns on that device.
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Address of this device's scratch (the calling thread's current device),
// looked up once per device.
unsigned char *scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    MHGP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned no address for the closed-ball scratch");
    g_scratch_ptr[device] = static_cast<unsigned char *>(p);
  }
  return g_scratch_ptr[device];
}

// Byte layout of the scratch for Q queries with `cap` interior ids each;
// every part starts on a 256-byte boundary.
struct NoAllocLayout {
  size_t ids, sizes, interior, nin, nsh, flags, und, nund, end;
  uint64_t und_cap, und_pow2;  // pairs recorded; room for the padded sort
  NoAllocLayout(int64_t Q, uint32_t cap) {
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    und_cap = std::max<uint64_t>(1024, (uint64_t)Q * 4);
    und_pow2 = 1;
    while (und_pow2 < und_cap) und_pow2 <<= 1;
    ids = 0;
    sizes = up(ids + 16 * (size_t)Q);
    interior = up(sizes + (size_t)Q);
    nin = up(interior + 4 * (size_t)Q * std::max<uint32_t>(cap, 1));
    nsh = up(nin + 4 * (size_t)Q);
    flags = up(nsh + 4 * (size_t)Q);
    und = up(flags + 4 * (size_t)Q);
    nund = up(und + 8 * (size_t)und_pow2);
    end = nund + 8;
  }
};

// Adds id to a row holding the `cap` smallest interior ids seen so far
// (`k` of them in use): appended while there is room, else it replaces the
// row's largest id when smaller.
void keep_smallest(uint32_t *row, uint32_t k, uint32_t cap, uint32_t id) {
  if (k < cap) {
    row[k] = id;
    return;
  }
  if (!cap) return;
  uint32_t *mx = std::max_element(row, row + cap);
  if (id < *mx) *mx = id;
}

}  // namespace

void closed_ball(Context &c, const uint32_t *ids, const uint8_t *sizes, int64_t Q, uint32_t cap,
                 uint32_t *interior, uint32_t *nin, uint32_t *nshell) {
  require_cloud(c);
  if (Q < 0 || (Q > 0 && (!ids || !sizes || !nin || !nshell)) || (cap > 0 && Q > 0 && !interior))
    throw std::invalid_argument("invalid closed-ball query arrays");
  if (Q == 0) return;
  for (int64_t q = 0; q < Q; ++q) {
    const int m = sizes[q];
    if (m < 1 || m > 4) throw std::invalid_argument("support sizes must lie in 1..4");
    for (int i = 0; i < m; ++i) {
      if (ids[4 * q + i] >= (uint64_t)c.n) throw std::invalid_argument("support PointId out of range");
      for (int j = 0; j < i; ++j)
        if (ids[4 * q + i] == ids[4 * q + j]) throw std::invalid_argument("repeated support PointId");
    }
  }
  hipStream_t s = c.stream;
  DevBuf<uint32_t> d_ids, d_int, d_nin, d_nsh, d_flags, d_und;
  DevBuf<uint8_t> d_sizes;
  DevBuf<uint64_t> d_nund;
  const uint64_t und_cap = std::max<uint64_t>(1024, (uint64_t)Q * 4);
  d_ids.ensure(4 * (size_t)Q);
  d_sizes.ensure(Q);
  d_int.ensure((size_t)Q * std::max<uint32_t>(cap, 1));
  d_nin.ensure(Q);
  d_nsh.ensure(Q);
  d_flags.ensure(Q);
  d_und.ensure(2 * und_cap);
  d_nund.ensure(1);
  MHGP_HIP(hipMemcpyAsync(d_ids.p, ids, 4 * (size_t)Q * 4, hipMemcpyHostToDevice, s));
  MHGP_HIP(hipMemcpyAsync(d_sizes.p, sizes, (size_t)Q, hipMemcpyHostToDevice, s));
  MHGP_HIP(hipMemsetAsync(d_nund.p, 0, 8, s));
  BallParams B{};
  B.ids = d_ids.p; B.sizes = d_sizes.p; B.queries = Q; B.cap = cap;
  B.X = c.X.p; B.Y = c.Y.p; B.Z = c.Z.p;
  B.nodes = c.nodes.p; B.LX = c.LX.p; B.LY = c.LY.p; B.LZ = c.LZ.p; B.leaf_id = c.leaf_id.p; B.n = c.n;
  B.interior = d_int.p; B.nin = d_nin.p; B.nshell = d_nsh.p; B.flags = d_flags.p;
  B.und = d_und.p; B.nund = (unsigned long long *)d_nund.p; B.und_cap = und_cap;
  if (c.n > 1) {
    hipLaunchKernelGGL(k_closed_ball, dim3(blocks_for(Q, 64)), dim3(64), 0, s, B);
    MHGP_LAUNCH_CHECK();
  } else {
    MHGP_HIP(hipMemsetAsync(d_nin.p, 0, Q * 4, s));
    MHGP_HIP(hipMemsetAsync(d_flags.p, 0, Q * 4, s));
    std::vector<uint32_t> ones(Q);
    for (int64_t q = 0; q < Q; ++q) ones[q] = sizes[q];
    MHGP_HIP(hipMemcpyAsync(d_nsh.p, ones.data(), Q * 4, hipMemcpyHostToDevice, s));
  }
  std::vector<uint32_t> flags(Q), inter((size_t)Q * cap);
  uint64_t nund = 0;
  MHGP_HIP(hipMemcpyAsync(nin, d_nin.p, Q * 4, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipMemcpyAsync(nshell, d_nsh.p, Q * 4, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipMemcpyAsync(flags.data(), d_flags.p, Q * 4, hipMemcpyDeviceToHost, s));
  if (cap) MHGP_HIP(hipMemcpyAsync(inter.data(), d_int.p, (size_t)Q * cap * 4, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipMemcpyAsync(&nund, d_nund.p, 8, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> und(2 * std::min(nund, und_cap));
  if (!und.empty()) copy_wait(und.data(), d_und.p, und.size() * 4, hipMemcpyDeviceToHost, s);
  // exact settlement: undecided points of device queries, then whole host queries
  const ExactCloud *ex = nullptr;
  auto exact = [&]() -> const ExactCloud & {
    if (!ex) ex = &exact_cloud(c);
    return *ex;
  };
  auto add_inside = [&](int64_t q, uint32_t id) {
    keep_smallest(inter.data() + (size_t)q * cap, std::min(nin[q], cap), cap, id);
    nin[q]++;
  };
  if (!und.empty()) {
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (size_t i = 0; i < und.size(); i += 2)
      if (!(flags[und[i]] & kHostQuery)) pairs.push_back({und[i], und[i + 1]});
    std::sort(pairs.begin(), pairs.end());
    ExactSupport S;
    int64_t last = -1;
    for (const auto &pq : pairs) {
      const int64_t q = pq.first;
      if (q != last) {
        int64_t sid[4];
        for (int i = 0; i < sizes[q]; ++i) sid[i] = ids[4 * q + i];
        if (!S.analyze(exact(), sid, sizes[q])) throw std::logic_error("a device query had a dependent support");
        last = q;
      }
      const int sd = S.side(exact(), pq.second);
      if (sd < 0) add_inside(q, pq.second);
      else if (sd == 0) nshell[q]++;
    }
  }
  for (int64_t q = 0; q < Q; ++q) {
    if (!(flags[q] & kHostQuery)) continue;
    nin[q] = 0;
    nshell[q] = sizes[q];
    int64_t sid[4];
    for (int i = 0; i < sizes[q]; ++i) sid[i] = ids[4 * q + i];
    ExactSupport S;
    if (!S.analyze(exact(), sid, sizes[q]))
      throw std::invalid_argument("the support of a closed-ball query is affinely dependent");
    for (int64_t x = 0; x < c.n; ++x) {
      bool on = false;
      for (int i = 0; i < sizes[q]; ++i) on |= sid[i] == x;
      if (on) continue;
      const int sd = S.side(exact(), x);
      if (sd < 0) add_inside(q, (uint32_t)x);
      else if (sd == 0) nshell[q]++;
    }
  }
  if (cap && interior)
    for (int64_t q = 0; q < Q; ++q) {
      const uint32_t k = std::min(nin[q], cap);
      std::sort(inter.begin() + (size_t)q * cap, inter.begin() + (size_t)q * cap + k);
      std::copy(inter.begin() + (size_t)q * cap, inter.begin() + (size_t)q * cap + k, interior + (size_t)q * cap);
      std::fill(interior + (size_t)q * cap + k, interior + (size_t)(q + 1) * cap, kInvalid);
    }
}

size_t closed_ball_scratch_bytes() { return kNoAllocScratchBytes; }

// Same results as closed_ball, with every device buffer taken from the
// file-scope scratch: no device allocation, no hipMalloc or hipFree.  The
// rare exact settlement of undecided points runs on the host, in host memory
// (the context's exact mirror of the cloud is built on its first use).
void closed_ball_noalloc(Context &c, const uint32_t *ids, const uint8_t *sizes, int64_t Q, uint32_t cap,
                         uint32_t *interior, uint32_t *nin, uint32_t *nshell) {
  require_cloud(c);
  if (Q < 0 || (Q > 0 && (!ids || !sizes || !nin || !nshell)) || (cap > 0 && Q > 0 && !interior))
    throw std::invalid_argument("invalid closed-ball query arrays");
  if (Q == 0) return;
  for (int64_t q = 0; q < Q; ++q) {
    const int m = sizes[q];
    if (m < 1 || m > 4) throw std::invalid_argument("support sizes must lie in 1..4");
    for (int i = 0; i < m; ++i) {
      if (ids[4 * q + i] >= (uint64_t)c.n) throw std::invalid_argument("support PointId out of range");
      for (int j = 0; j < i; ++j)
        if (ids[4 * q + i] == ids[4 * q + j]) throw std::invalid_argument("repeated support PointId");
    }
  }
  const NoAllocLayout L(Q, cap);
  if (L.end > kNoAllocScratchBytes)
    throw CapacityError("closed-ball queries need " + std::to_string(L.end) + " bytes of device scratch; the " +
                        "allocation-free entry point has " + std::to_string(kNoAllocScratchBytes) +
                        " (fewer queries or a smaller interior cap per call)");
  unsigned char *base = scratch_ptr(c.device);
  uint32_t *d_ids = reinterpret_cast<uint32_t *>(base + L.ids);
  uint8_t *d_sizes = base + L.sizes;
  uint32_t *d_int = reinterpret_cast<uint32_t *>(base + L.interior);
  uint32_t *d_nin = reinterpret_cast<uint32_t *>(base + L.nin);
  uint32_t *d_nsh = reinterpret_cast<uint32_t *>(base + L.nsh);
  uint32_t *d_flags = reinterpret_cast<uint32_t *>(base + L.flags);
  uint32_t *d_und = reinterpret_cast<uint32_t *>(base + L.und);
  unsigned long long *d_nund = reinterpret_cast<unsigned long long *>(base + L.nund);
  hipStream_t s = c.stream;
  constexpr uint32_t kFlag = 0x80000000u;
  auto add_inside = [&](int64_t q, uint32_t id) {
    keep_smallest(interior + (size_t)q * cap, std::min(nin[q] & ~kFlag, cap), cap, id);
    nin[q]++;
  };
  if (c.n > 1) {
    MHGP_HIP(hipMemcpyAsync(d_ids, ids, 16 * (size_t)Q, hipMemcpyHostToDevice, s));
    MHGP_HIP(hipMemcpyAsync(d_sizes, sizes, (size_t)Q, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, d_nund);
    MHGP_LAUNCH_CHECK();
    BallParams B{};
    B.ids = d_ids; B.sizes = d_sizes; B.queries = Q; B.cap = cap;
    B.X = c.X.p; B.Y = c.Y.p; B.Z = c.Z.p;
    B.nodes = c.nodes.p; B.LX = c.LX.p; B.LY = c.LY.p; B.LZ = c.LZ.p; B.leaf_id = c.leaf_id.p; B.n = c.n;
    B.interior = d_int; B.nin = d_nin; B.nshell = d_nsh; B.flags = d_flags;
    B.und = d_und; B.nund = d_nund; B.und_cap = L.und_cap;
    hipLaunchKernelGGL(k_closed_ball, dim3(blocks_for(Q, 64)), dim3(64), 0, s, B);
    MHGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_fold_flags, dim3(blocks_for(Q, 256)), dim3(256), 0, s, d_nin, d_flags, Q);
    MHGP_LAUNCH_CHECK();
    uint64_t nund = 0;
    MHGP_HIP(hipMemcpyAsync(nin, d_nin, 4 * (size_t)Q, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(nshell, d_nsh, 4 * (size_t)Q, hipMemcpyDeviceToHost, s));
    if (cap) MHGP_HIP(hipMemcpyAsync(interior, d_int, 4 * (size_t)Q * cap, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(&nund, d_nund, 8, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipStreamSynchronize(s));
    nund = std::min<uint64_t>(nund, L.und_cap);
    // settle the pairs in (query, point) order, as mhgp_closed_ball does: it
    // keeps, of a query's undecided inside points, the smallest ids when the
    // interior list overflows.  More pairs than one host chunk are sorted on
    // the device, in the scratch, by a bitonic network.
    constexpr uint64_t kChunk = 1024;
    const bool device_sorted = nund > kChunk;
    if (device_sorted) {
      uint64_t padded = 1;
      while (padded < nund) padded <<= 1;
      uint64_t *keys = reinterpret_cast<uint64_t *>(d_und);
      hipLaunchKernelGGL(k_und_keys, dim3(blocks_for((int64_t)padded, 256)), dim3(256), 0, s, keys, nund, padded);
      for (uint64_t k = 2; k <= padded; k <<= 1)
        for (uint64_t j = k >> 1; j > 0; j >>= 1)
          hipLaunchKernelGGL(k_bitonic_stage, dim3(blocks_for((int64_t)padded, 256)), dim3(256), 0, s, keys, padded,
                             j, k);
      MHGP_LAUNCH_CHECK();
    }
    // undecided points of the device queries, settled exactly in chunks
    // copied through a fixed host array
    uint64_t pairs[kChunk];
    ExactSupport S;
    int64_t last = -1;
    for (uint64_t at = 0; at < nund; at += kChunk) {
      const uint64_t cnt = std::min(kChunk, nund - at);
      copy_wait(pairs, d_und + 2 * at, 8 * cnt, hipMemcpyDeviceToHost, s);
      if (!device_sorted) {
        // one chunk holds every pair: (query, point) keys sorted here
        for (uint64_t i = 0; i < cnt; ++i) {
          const uint32_t *w = reinterpret_cast<const uint32_t *>(&pairs[i]);
          pairs[i] = ((uint64_t)w[0] << 32) | w[1];
        }
        std::sort(pairs, pairs + cnt);
      }
      for (uint64_t i = 0; i < cnt; ++i) {
        const int64_t q = (int64_t)(pairs[i] >> 32);
        const uint32_t x = (uint32_t)pairs[i];
        if (nin[q] & kFlag) continue;
        if (q != last) {
          int64_t sid[4];
          for (int k = 0; k < sizes[q]; ++k) sid[k] = ids[4 * q + k];
          if (!S.analyze(exact_cloud(c), sid, sizes[q])) throw std::logic_error("a device query had a dependent support");
          last = q;
        }
        const int sd = S.side(exact_cloud(c), x);
        if (sd < 0) add_inside(q, x);
        else if (sd == 0) nshell[q]++;
      }
    }
  } else {
    // one point: every query is that point alone, on its own zero sphere
    for (int64_t q = 0; q < Q; ++q) { nin[q] = 0; nshell[q] = sizes[q]; }
  }
  // queries the device handed over whole
  for (int64_t q = 0; q < Q; ++q) {
    if (!(nin[q] & kFlag)) continue;
    nin[q] = kFlag;
    nshell[q] = sizes[q];
    int64_t sid[4];
    for (int i = 0; i < sizes[q]; ++i) sid[i] = ids[4 * q + i];
    ExactSupport S;
    if (!S.analyze(exact_cloud(c), sid, sizes[q]))
      throw std::invalid_argument("the support of a closed-ball query is affinely dependent");
    for (int64_t x = 0; x < c.n; ++x) {
      bool on = false;
      for (int i = 0; i < sizes[q]; ++i) on |= sid[i] == x;
      if (on) continue;
      const int sd = S.side(exact_cloud(c), x);
      if (sd < 0) add_inside(q, (uint32_t)x);
      else if (sd == 0) nshell[q]++;
    }
  }
  for (int64_t q = 0; q < Q; ++q) {
    nin[q] &= ~kFlag;
    if (!cap) continue;
    const uint32_t k = std::min(nin[q], cap);
    std::sort(interior + (size_t)q * cap, interior + (size_t)q * cap + k);
    std::fill(interior + (size_t)q * cap + k, interior + (size_t)(q + 1) * cap, kInvalid);
  }
}

}  // namespace mhgp
