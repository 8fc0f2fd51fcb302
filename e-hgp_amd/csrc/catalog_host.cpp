// Host side of the critical-sphere catalogue: exact settlement of the work
// the device left undecided; the canonical order is built on the device.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "pipeline.hpp"
#include "supports.hpp"
#include "timing.hpp"
#include "tower_gpu.hpp"
#include "tower_state.hpp"

namespace mhgp {

void supports_device(Context &c, int k_max, std::vector<EventRec> &events,
                     std::vector<FallbackRec> &fbs, std::vector<uint32_t> &fb_region,
                     mhgp_catalog_info *info);

namespace {

struct ExactJob {
  const ExactCloud &ex;
  int s_max;
  int scale_exp;
  std::vector<EventRec> &out;
  int64_t predicates = 0;

  // Exact classification of a support against a region (the region holds
  // every point of any relevant sphere through the anchor).
  void support(const int64_t *ids, int m, const uint32_t *region, uint32_t len) {
    ExactSupport S;
    predicates++;
    if (!S.analyze(ex, ids, m) || !S.well) return;
    uint32_t interior[kMaxInterior];
    int nint = 0, extra = 0;
    int64_t witness = -1;  // least extra-shell point
    for (uint32_t s = 0; s < len; ++s) {
      int64_t x = region[s];
      bool on = false;
      for (int k = 0; k < m; ++k) on |= (x == ids[k]);
      if (on) continue;
      int side = S.side(ex, x);
      predicates++;
      if (side < 0) {
        if (nint < kMaxInterior) interior[nint] = (uint32_t)x;
        nint++;
        if (nint + m > s_max) return;
      } else if (side == 0) {
        extra++;
        if (witness < 0 || x < witness) witness = x;
      }
    }
    EventRec e;
    for (int k = 0; k < 4; ++k) e.ids[k] = k < m ? (uint32_t)ids[k] : kInvalid;
    std::sort(interior, interior + nint);
    e.meta = (uint32_t)m | ((uint32_t)(nint + m) << 8) | ((uint32_t)nint << 16) |
             ((extra ? 1u : 0u) << 24) | ((uint32_t)std::min(extra, 127) << 25);
    for (int k = 0; k < kMaxInterior; ++k) e.interior[k] = k < nint ? interior[k] : kInvalid;
    // an extra-shell record keeps its least extra-shell point in the last
    // interior slot, which a support of 2..4 points never fills (nint <= 9)
    if (extra) e.interior[kMaxInterior - 1] = (uint32_t)witness;
    // in the device's frame: coordinates there are scaled by 2^scale_exp
    e.level = ratio_to_double(S.lnum, S.lden, 2 * (ex.e0() + scale_exp));
    e.level_err = 1e-12f;
    out.push_back(e);
  }
};

}  // namespace

void supports(Context &c, int k_max, mhgp_catalog_info *info) {
  if (k_max < 1 || k_max > 10) throw std::invalid_argument("k_max must lie in 1..10");
  *info = mhgp_catalog_info{};
  if (c.tower) c.tower->valid = false;
  std::vector<EventRec> unused;
  std::vector<FallbackRec> fbs;
  std::vector<uint32_t> region;
  supports_device(c, k_max, unused, fbs, region, info);
  StageTimer tm(c.stream);
  CatalogBuffers &B = *c.catalog;
  const int smax = B.s_max;
  int64_t diagnostics = 0;
  if (!fbs.empty()) {
    std::vector<EventRec> settled;
    const ExactCloud &ex = exact_cloud(c);
    ExactJob job{ex, smax, c.scale_exp, settled};
    for (const FallbackRec &f : fbs) {
      if (f.region_off == kInvalid) throw std::runtime_error("fallback without a published region");
      // device records index host arrays: refuse any that leaves them
      if (f.kind < 2 || f.kind > 4 || (uint64_t)f.region_off + f.region_len > region.size())
        throw std::runtime_error("malformed fallback record from the device");
      for (uint32_t q = 0; q < (f.kind == 3 ? 3u : f.kind); ++q)
        if (f.ids[q] >= (uint64_t)c.n) throw std::runtime_error("malformed fallback record from the device");
      const uint32_t *reg = region.data() + f.region_off;
      const uint32_t len = f.region_len;
      for (uint32_t q = 0; q < len; ++q)
        if (reg[q] >= (uint64_t)c.n) throw std::runtime_error("malformed fallback region from the device");
      if (f.kind == 2) {
        int64_t ids[2] = {f.ids[0], f.ids[1]};
        job.support(ids, 2, reg, len);
      } else if (f.kind == 4) {
        int64_t ids[4] = {f.ids[0], f.ids[1], f.ids[2], f.ids[3]};
        job.support(ids, 4, reg, len);
      } else if (f.kind == 3) {
        int64_t ids[4] = {f.ids[0], f.ids[1], f.ids[2], 0};
        job.support(ids, 3, reg, len);
        for (uint32_t s = 0; s < len; ++s) {
          if (reg[s] <= f.ids[2]) continue;
          ids[3] = reg[s];
          job.support(ids, 4, reg, len);
        }
      }
    }
    info->fallback_count = job.predicates;
    // append the settled records behind the device's
    if (!settled.empty()) {
      const int64_t total = B.count + (int64_t)settled.size();
      if ((uint64_t)total > B.events.cap) {
        DevBuf<EventRec> grown;
        grown.ensure(total);
        copy_wait(grown.p, B.events.p, B.count * sizeof(EventRec), hipMemcpyDeviceToDevice, c.stream);
        std::swap(grown.p, B.events.p);
        std::swap(grown.cap, B.events.cap);
      }
      copy_wait(B.events.p + B.count, settled.data(), settled.size() * sizeof(EventRec),
                         hipMemcpyHostToDevice, c.stream);
      B.count = total;
      for (const EventRec &e : settled) diagnostics += (e.meta >> 24) & 1;
    }
  }
  tm.mark("catalogue host fallback");
  // the tower's neighbour lists run beside the level order
  if (B.count > 0) tower_prefetch_neighbours(c);
  order_catalogue(c, B);
  info->event_count = B.count - diagnostics;
  info->diagnostic_count = diagnostics;
  info->s_max = smax;
  // status (conditional when anchors were left open) comes from the device pass
  B.info = *info;
  B.valid = true;
}

void supports_fetch(Context &c, uint32_t *ids, uint32_t *meta, uint32_t *interior, double *level,
                    int64_t *level_class) {
  if (!c.catalog || !c.catalog->valid) throw StateError("no catalogue has been built");
  CatalogBuffers &B = *c.catalog;
  const int64_t n = B.count;
  std::vector<EventRec> ev(n);
  std::vector<uint32_t> order(n), cls(n);
  if (n) {
    copy_wait(ev.data(), B.events.p, n * sizeof(EventRec), hipMemcpyDeviceToHost, c.stream);
    copy_wait(order.data(), B.order.p, n * 4, hipMemcpyDeviceToHost, c.stream);
    copy_wait(cls.data(), B.cls.p, n * 4, hipMemcpyDeviceToHost, c.stream);
  }
  const int width = std::max(B.s_max - 1, 1);
  for (int64_t i = 0; i < n; ++i) {
    const EventRec &e = ev[order[i]];
    if (ids)
      for (int k = 0; k < 4; ++k) ids[4 * i + k] = e.ids[k];
    if (meta) meta[i] = e.meta;
    const int nint = (int)((e.meta >> 16) & 0xFFu);
    if (interior)
      for (int k = 0; k < width; ++k) interior[width * i + k] = k < nint ? e.interior[k] : kInvalid;
    if (level) level[i] = std::ldexp(e.level, -2 * c.scale_exp);
    if (level_class) level_class[i] = cls[i];
  }
}

}  // namespace mhgp

namespace mhgp {

void supports_radii(Context &c, double *radius_out) {
  if (!c.catalog || !c.catalog->valid) throw StateError("no catalogue has been built");
  if (!radius_out) throw std::invalid_argument("null output buffer");
  const int64_t n = c.n;
  if (n < 2 || c.catalog->s_max < 2) {
    for (int64_t i = 0; i < n; ++i) radius_out[i] = 0.0;
    return;
  }
  copy_wait(radius_out, c.catalog->radius.p, n * 8, hipMemcpyDeviceToHost, c.stream);
  // back to the input unit (an exact power of two); rounded up past the ulp
  // the scaling cannot lose
  for (int64_t i = 0; i < n; ++i) radius_out[i] = std::ldexp(radius_out[i], -c.scale_exp);
}

}  // namespace mhgp

namespace mhgp {

void supports_diagnostics(Context &c, uint32_t *shell_count, int64_t *witness, int64_t *exterior_count) {
  if (!c.catalog || !c.catalog->valid) throw StateError("no catalogue has been built");
  CatalogBuffers &B = *c.catalog;
  const int64_t n = B.count;
  std::vector<EventRec> ev(n);
  std::vector<uint32_t> order(n);
  if (n) {
    copy_wait(ev.data(), B.events.p, n * sizeof(EventRec), hipMemcpyDeviceToHost, c.stream);
    copy_wait(order.data(), B.order.p, n * 4, hipMemcpyDeviceToHost, c.stream);
  }
  for (int64_t i = 0; i < n; ++i) {
    const EventRec &e = ev[order[i]];
    const int64_t m = e.meta & 0xFFu, rank = (e.meta >> 8) & 0xFFu;
    const bool deg = (e.meta >> 24) & 1u;
    int64_t extra = deg ? (int64_t)(e.meta >> 25) : 0;
    if (extra >= 127) {
      // the count saturates in the record: recount the sphere exactly
      const ExactCloud &exc = exact_cloud(c);
      int64_t ids[4];
      for (int k = 0; k < m; ++k) ids[k] = e.ids[k];
      ExactSupport S;
      if (!S.analyze(exc, ids, (int)m)) throw std::logic_error("a catalogue support is affinely dependent");
      extra = 0;
      for (int64_t x = 0; x < c.n; ++x) {
        bool on = false;
        for (int k = 0; k < m; ++k) on |= ids[k] == x;
        if (!on && S.side(exc, x) == 0) extra++;
      }
    }
    if (shell_count) shell_count[i] = (uint32_t)(m + extra);
    if (witness) witness[i] = deg ? (int64_t)e.interior[kMaxInterior - 1] : -1;
    if (exterior_count) exterior_count[i] = c.n - rank - extra;
  }
}

}  // namespace mhgp
