// Canonical order of the catalogue on the device.
//
// Records are radix-sorted by the bits of their FP64 squared radius.  Two
// neighbours whose FP64 intervals are disjoint are provably ordered and start
// a new exact level class.  Each run whose order the FP64 intervals leave
// open is re-evaluated in double-double (level_dd.hpp), sorted and certified
// by one device thread; only runs that stay open (exact ties) are gathered,
// ordered exactly on the host by (exact level, rank, support ids) and
// scattered back.  A device scan turns the class breaks into dense classes.
#include <algorithm>
#include <cstdio>
#include <exception>
#include <numeric>
#include <thread>

#include "pipeline.hpp"
#include "scan.hpp"
#include "level_dd.hpp"
#include "level_links.hpp"
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

// Upper bound of every record's level interval, in sorted order (the bits of
// a non-negative double order like the double).
__global__ void k_level_upper(const EventRec *__restrict__ ev, const uint32_t *__restrict__ order, int64_t n,
                              uint64_t *hi) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const EventRec &a = ev[order[i]];
  hi[i] = dbits(level_upper(a.level, a.level_err));
}

// prefmax: inclusive prefix maximum of the upper bounds (level_links.hpp).
__global__ void k_links(const EventRec *__restrict__ ev, const uint32_t *__restrict__ order,
                        const uint64_t *__restrict__ prefmax, int64_t n, uint32_t *link) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0) { link[0] = 1; return; }
  const EventRec &b = ev[order[i]];
  link[i] = level_link(level_lower(b.level, b.level_err), __longlong_as_double((long long)prefmax[i - 1])) ? 1u : 0u;
}

// Start of every run of records whose mutual order the FP64 intervals leave open.
// Each block takes kRunTile consecutive records and appends its starts with
// one global atomic (runs are independent, their list order is free).
constexpr int kRunTile = 16 * 256;
__global__ void __launch_bounds__(256) k_run_starts(const uint32_t *__restrict__ link, int64_t n, uint32_t *starts,
                                                    unsigned long long *count, uint64_t cap) {
  __shared__ uint32_t local[kRunTile];
  __shared__ uint32_t nlocal;
  __shared__ unsigned long long base;
  if (threadIdx.x == 0) nlocal = 0;
  __syncthreads();
  const int64_t tile = (int64_t)blockIdx.x * kRunTile;
  for (int j = threadIdx.x; j < kRunTile; j += blockDim.x) {
    const int64_t i = tile + j;
    if (i + 1 < n && link[i] && !link[i + 1]) local[atomicAdd(&nlocal, 1u)] = (uint32_t)i;
  }
  __syncthreads();
  if (threadIdx.x == 0) base = nlocal ? atomicAdd(count, (unsigned long long)nlocal) : 0ull;
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < nlocal; j += blockDim.x)
    if (base + j < cap) starts[base + j] = local[j];
}

constexpr int kRunMax = 32;

__device__ inline bool canonical_less(const DDE &lx, const EventRec &x, const DDE &ly, const EventRec &y) {
  if (dd_less(lx, ly)) return true;
  if (dd_less(ly, lx)) return false;
  const uint32_t rx = (x.meta >> 8) & 0xff, ry = (y.meta >> 8) & 0xff;
  if (rx != ry) return rx < ry;
  for (int q = 0; q < 4; ++q)
    if (x.ids[q] != y.ids[q]) return x.ids[q] < y.ids[q];
  return false;
}

// One thread per open run: double-double levels, sort, certify every link.
// Runs that stay open (exact ties, or longer than kRunMax) go to the host.
__global__ void k_resolve_runs(const EventRec *__restrict__ ev, uint32_t *order, uint32_t *link, int64_t n,
                               const double *__restrict__ X, const double *__restrict__ Y,
                               const double *__restrict__ Z, const uint32_t *__restrict__ starts,
                               uint32_t count, uint32_t *host_start, uint32_t *host_len,
                               unsigned long long *host_count) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= count) return;
  const int64_t lo = starts[r];
  int64_t hi = lo + 1;
  while (hi < n && !link[hi]) hi++;
  const int64_t len = hi - lo;
  bool open = len > kRunMax;
  if (!open) {
    uint32_t idx[kRunMax];
    DDE lv[kRunMax];
    for (int k = 0; k < len; ++k) {
      idx[k] = order[lo + k];
      const EventRec &e = ev[idx[k]];
      lv[k] = dd_level(X, Y, Z, e.ids, e.meta & 0xff);
    }
    for (int k = 1; k < len; ++k) {  // insertion sort in canonical order
      const uint32_t vi = idx[k];
      const DDE vl = lv[k];
      int j = k - 1;
      while (j >= 0 && canonical_less(vl, ev[vi], lv[j], ev[idx[j]])) {
        idx[j + 1] = idx[j];
        lv[j + 1] = lv[j];
        j--;
      }
      idx[j + 1] = vi;
      lv[j + 1] = vl;
    }
    for (int k = 1; k < len && !open; ++k) open = dd_order(lv[k - 1], lv[k]) != 1;
    if (!open) {
      for (int k = 0; k < len; ++k) order[lo + k] = idx[k];
      for (int k = 1; k < len; ++k) link[lo + k] = 1;
      return;
    }
  }
  const unsigned long long s = atomicAdd(host_count, 1ull);
  host_start[s] = (uint32_t)lo;
  host_len[s] = (uint32_t)len;
}

__global__ void k_gather_runs(const EventRec *__restrict__ ev, const uint32_t *__restrict__ order,
                              const uint32_t *__restrict__ pos, int64_t count, uint32_t *gidx,
                              EventRec *grec) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const uint32_t i = order[pos[t]];
  gidx[t] = i;
  grec[t] = ev[i];
}

__global__ void k_scatter_runs(const uint32_t *__restrict__ pos, const uint32_t *__restrict__ new_idx,
                               const uint32_t *__restrict__ new_link, int64_t count, uint32_t *order,
                               uint32_t *link) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  order[pos[t]] = new_idx[t];
  link[pos[t]] = new_link[t];
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

// Exact canonical order of one run of host records [b, e) of the gathered
// arrays: writes the permuted record indices and the class breaks.  Two
// records whose FP64 level intervals (level_lower .. level_upper) are
// disjoint compare by them; the exact rational levels
// are made only for the pairs whose intervals overlap (once per record).
void exact_run(const ExactCloud &ex, const uint32_t *gidx, const EventRec *grec, int64_t b, int64_t e,
               uint32_t *new_idx, uint32_t *new_link) {
  const int64_t len = e - b;
  std::vector<BigInt> num(len), den(len);
  std::vector<char> made(len, 0);
  std::vector<double> lo(len), hi(len);
  for (int64_t k = 0; k < len; ++k) {
    lo[k] = level_lower(grec[b + k].level, grec[b + k].level_err);
    hi[k] = level_upper(grec[b + k].level, grec[b + k].level_err);
  }
  auto level_cmp = [&](int64_t x, int64_t y) {
    if (hi[x] < lo[y]) return -1;
    if (hi[y] < lo[x]) return 1;
    for (int64_t z : {x, y})
      if (!made[z]) {
        exact_level_of(ex, grec[b + z], num[z], den[z]);
        made[z] = 1;
      }
    return compare_levels(num[x], den[x], num[y], den[y]);
  };
  std::vector<int64_t> perm(len);
  std::iota(perm.begin(), perm.end(), 0);
  std::sort(perm.begin(), perm.end(), [&](int64_t x, int64_t y) {
    int cmp = level_cmp(x, y);
    if (cmp) return cmp < 0;
    const EventRec &rx = grec[b + x], &ry = grec[b + y];
    uint32_t kx = (rx.meta >> 8) & 0xff, ky = (ry.meta >> 8) & 0xff;
    if (kx != ky) return kx < ky;
    for (int q = 0; q < 4; ++q)
      if (rx.ids[q] != ry.ids[q]) return rx.ids[q] < ry.ids[q];
    return false;
  });
  for (int64_t k = 0; k < len; ++k) {
    new_idx[b + k] = gidx[b + perm[k]];
    // the link into the run was certain; inside it, a break at every exact increase
    new_link[b + k] = k == 0 ? 1u : (level_cmp(perm[k - 1], perm[k]) != 0);
  }
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
  // the sort keys are free once sorted: they take the prefix maxima
  hipLaunchKernelGGL(k_level_upper, dim3(blocks_for(n, 256)), dim3(256), 0, s, B.events.p, B.order.p, n,
                     B.sort_keys.p);
  c.scanner.inclusive_max(B.sort_keys.p, n, s);
  hipLaunchKernelGGL(k_links, dim3(blocks_for(n, 256)), dim3(256), 0, s, B.events.p, B.order.p, B.sort_keys.p, n,
                     B.cls.p);
  // open runs: starts, then a double-double pass over each
  const uint64_t cap = std::max<int64_t>(1 << 16, n / 4);
  B.ambiguous.ensure(3 * cap);
  uint32_t *starts = B.ambiguous.p, *host_start = starts + cap, *host_len = host_start + cap;
  DevBuf<uint64_t> &cnt = B.run_count;
  cnt.ensure(2);
  MHGP_HIP(hipMemsetAsync(cnt.p, 0, 16, s));
  hipLaunchKernelGGL(k_run_starts, dim3(blocks_for(n, kRunTile)), dim3(256), 0, s, B.cls.p, n, starts,
                     (unsigned long long *)cnt.p, cap);
  uint64_t runs = 0;
  copy_wait(&runs, cnt.p, 8, hipMemcpyDeviceToHost, s);
  if (runs > cap) throw std::runtime_error("too many open level runs");
  uint64_t nh = 0;
  if (runs) {
    hipLaunchKernelGGL(k_resolve_runs, dim3(blocks_for(runs, 64)), dim3(64), 0, s, B.events.p, B.order.p,
                       B.cls.p, n, c.X.p, c.Y.p, c.Z.p, starts, (uint32_t)runs, host_start, host_len,
                       (unsigned long long *)cnt.p + 1);
    MHGP_LAUNCH_CHECK();
    copy_wait(&nh, cnt.p + 1, 8, hipMemcpyDeviceToHost, s);
  }
  tm.mark("catalogue level runs (device)");
  B.open_runs = (int64_t)runs;
  B.host_runs = (int64_t)nh;
  if (nh) {
    // runs the double-double pass left open: exact order on the host, one
    // gather and one scatter for all of them
    std::vector<uint32_t> hs(nh), hl(nh);
    MHGP_HIP(hipMemcpyAsync(hs.data(), host_start, nh * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(hl.data(), host_len, nh * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> off(nh + 1, 0);
    for (uint64_t r = 0; r < nh; ++r) off[r + 1] = off[r] + hl[r];
    const int64_t T = off[nh];
    std::vector<uint32_t> pos(T);
    for (uint64_t r = 0; r < nh; ++r)
      for (uint32_t k = 0; k < hl[r]; ++k) pos[off[r] + k] = hs[r] + k;
    DevBuf<uint32_t> &dpos = B.run_pos, &dgidx = B.run_gidx, &dnew = B.run_new, &dlink = B.run_link;
    DevBuf<EventRec> &drec = B.run_rec;
    dpos.ensure(T); dgidx.ensure(T); dnew.ensure(T); dlink.ensure(T); drec.ensure(T);
    MHGP_HIP(hipMemcpyAsync(dpos.p, pos.data(), T * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gather_runs, dim3(blocks_for(T, 256)), dim3(256), 0, s, B.events.p, B.order.p, dpos.p,
                       T, dgidx.p, drec.p);
    std::vector<uint32_t> gidx(T), new_idx(T), new_link(T);
    std::vector<EventRec> grec(T);
    MHGP_HIP(hipMemcpyAsync(gidx.data(), dgidx.p, T * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(grec.data(), drec.p, T * sizeof(EventRec), hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipStreamSynchronize(s));
    if (tm.on) std::fprintf(stderr, "[mhgp] host runs: %lld records\n", (long long)T);
    tm.mark("catalogue runs gathered");
    const ExactCloud &ex = exact_cloud(c);
    tm.mark("catalogue exact cloud");
    const int threads = (int)std::max<int64_t>(1, std::min<int64_t>({16, (int64_t)std::thread::hardware_concurrency(),
                                                                      (int64_t)nh}));
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> errs(threads);
    for (int t = 0; t < threads; ++t)
      pool.emplace_back([&, t] {
        try {
          for (uint64_t r = t; r < nh; r += threads)
            exact_run(ex, gidx.data(), grec.data(), off[r], off[r + 1], new_idx.data(), new_link.data());
        } catch (...) {
          errs[t] = std::current_exception();
        }
      });
    for (auto &th : pool) th.join();
    for (auto &e : errs)
      if (e) std::rethrow_exception(e);
    tm.mark("catalogue runs ordered");
    MHGP_HIP(hipMemcpyAsync(dnew.p, new_idx.data(), T * 4, hipMemcpyHostToDevice, s));
    MHGP_HIP(hipMemcpyAsync(dlink.p, new_link.data(), T * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scatter_runs, dim3(blocks_for(T, 256)), dim3(256), 0, s, dpos.p, dnew.p, dlink.p, T,
                       B.order.p, B.cls.p);
    MHGP_HIP(hipStreamSynchronize(s));
  }
  if (tm.on)
    std::fprintf(stderr, "[mhgp] level runs: %llu open, %llu exact on the host\n", (unsigned long long)runs,
                 (unsigned long long)nh);
  tm.mark("catalogue level runs (host exact)");
  c.scanner.inclusive(B.cls.p, n, s);
  hipLaunchKernelGGL(k_class_from_scan, dim3(blocks_for(n, 256)), dim3(256), 0, s, B.cls.p, n);
  MHGP_LAUNCH_CHECK();
  tm.mark("catalogue level classes");
}

}  // namespace mhgp
