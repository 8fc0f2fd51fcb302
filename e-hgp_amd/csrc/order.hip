// Canonical order of the catalogue on the device.
//
// Records are radix-sorted by the bits of their FP64 squared radius.  Two
// neighbours whose FP64 intervals are disjoint are provably ordered and start
// a new exact level class; the short runs whose order is not provable are
// downloaded, ordered exactly on the host by (exact level, rank, support ids)
// and patched back.  A device scan turns the class breaks into dense classes.
#include <algorithm>
#include <numeric>

#include "pipeline.hpp"
#include "scan.hpp"
#include "supports.hpp"
#include "timing.hpp"

namespace mhgp {

namespace {

__global__ void k_level_keys(const EventRec *__restrict__ ev, int64_t n, uint64_t *keys,
                             uint32_t *vals) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = dbits(ev[i].level);
  vals[i] = (uint32_t)i;
}

__global__ void k_links(const EventRec *__restrict__ ev, const uint32_t *__restrict__ order, int64_t n,
                        uint32_t *link, uint32_t *amb, unsigned long long *namb, uint64_t cap) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0) { link[0] = 1; return; }
  const EventRec &a = ev[order[i - 1]], &b = ev[order[i]];
  const double ahi = a.level * (1.0 + 2.0 * (double)a.level_err);
  const double blo = b.level * (1.0 - 2.0 * (double)b.level_err);
  const bool certain = blo > ahi;
  link[i] = certain ? 1u : 0u;
  if (!certain) {
    unsigned long long s = atomicAdd(namb, 1ull);
    if (s < cap) amb[s] = (uint32_t)i;
  }
}

__global__ void k_class_from_scan(uint32_t *cls, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cls[i] -= 1;
}

void exact_level_of(const ExactCloud &ex, const EventRec &e, BigInt &num, BigInt &den) {
  int64_t ids[4];
  const int m = e.meta & 0xff;
  for (int k = 0; k < m; ++k) ids[k] = e.ids[k];
  ExactSupport S;
  if (!S.analyze(ex, ids, m)) throw std::runtime_error("dependent support in the catalogue");
  num = S.lnum;
  den = S.lden;
}

}  // namespace

void order_catalogue(Context &c, CatalogBuffers &B) {
  const int64_t n = B.count;
  hipStream_t s = c.stream;
  StageTimer tm(s);
  B.order.ensure(n);
  B.cls.ensure(n);
  if (n == 0) return;
  B.sort_keys.ensure(n);
  hipLaunchKernelGGL(k_level_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, B.events.p, n,
                     B.sort_keys.p, B.order.p);
  c.sorter.sort(B.sort_keys.p, B.order.p, n, 0, 64, s);
  tm.mark("catalogue level sort");
  const uint64_t cap = std::max<int64_t>(1 << 16, n / 16);
  B.ambiguous.ensure(cap);
  DevBuf<uint64_t> namb;
  namb.ensure(1);
  MHGP_HIP(hipMemsetAsync(namb.p, 0, 8, s));
  hipLaunchKernelGGL(k_links, dim3(blocks_for(n, 256)), dim3(256), 0, s, B.events.p, B.order.p, n,
                     B.cls.p, B.ambiguous.p, (unsigned long long *)namb.p, cap);
  uint64_t na = 0;
  MHGP_HIP(hipMemcpyAsync(&na, namb.p, 8, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipStreamSynchronize(s));
  if (na > cap) throw std::runtime_error("too many ambiguous level neighbours");
  if (na) {
    std::vector<uint32_t> pos(na);
    copy_wait(pos.data(), B.ambiguous.p, na * 4, hipMemcpyDeviceToHost, s);
    std::sort(pos.begin(), pos.end());
    const ExactCloud &ex = exact_cloud(c);
    for (size_t a = 0; a < pos.size();) {
      size_t b = a;
      while (b + 1 < pos.size() && pos[b + 1] == pos[b] + 1) b++;
      // run of sorted positions [lo, hi] with unproven internal order
      const int64_t lo = pos[a] - 1, hi = pos[b];
      const int64_t len = hi - lo + 1;
      std::vector<uint32_t> idx(len);
      copy_wait(idx.data(), B.order.p + lo, len * 4, hipMemcpyDeviceToHost, s);
      std::vector<EventRec> recs(len);
      for (int64_t k = 0; k < len; ++k)
        copy_wait(&recs[k], B.events.p + idx[k], sizeof(EventRec), hipMemcpyDeviceToHost, s);
      std::vector<BigInt> num(len), den(len);
      for (int64_t k = 0; k < len; ++k) exact_level_of(ex, recs[k], num[k], den[k]);
      std::vector<int64_t> perm(len);
      std::iota(perm.begin(), perm.end(), 0);
      std::sort(perm.begin(), perm.end(), [&](int64_t x, int64_t y) {
        int cmp = compare_levels(num[x], den[x], num[y], den[y]);
        if (cmp) return cmp < 0;
        uint32_t rx = (recs[x].meta >> 8) & 0xff, ry = (recs[y].meta >> 8) & 0xff;
        if (rx != ry) return rx < ry;
        for (int q = 0; q < 4; ++q)
          if (recs[x].ids[q] != recs[y].ids[q]) return recs[x].ids[q] < recs[y].ids[q];
        return false;
      });
      std::vector<uint32_t> new_idx(len), links(len);
      for (int64_t k = 0; k < len; ++k) {
        new_idx[k] = idx[perm[k]];
        links[k] = (k == 0) ? 1u
                            : (compare_levels(num[perm[k - 1]], den[perm[k - 1]], num[perm[k]], den[perm[k]]) != 0);
      }
      if (lo == 0) links[0] = 1;
      else {
        uint32_t l0;
        copy_wait(&l0, B.cls.p + lo, 4, hipMemcpyDeviceToHost, s);
        links[0] = l0;  // the link into the run was provable
      }
      copy_wait(B.order.p + lo, new_idx.data(), len * 4, hipMemcpyHostToDevice, s);
      copy_wait(B.cls.p + lo, links.data(), len * 4, hipMemcpyHostToDevice, s);
      a = b + 1;
    }
  }
  tm.mark("catalogue exact ties");
  Scanner sc;
  sc.inclusive(B.cls.p, n, s);
  hipLaunchKernelGGL(k_class_from_scan, dim3(blocks_for(n, 256)), dim3(256), 0, s, B.cls.p, n);
  MHGP_LAUNCH_CHECK();
  tm.mark("catalogue level classes");
}

}  // namespace mhgp
