// Host side of the critical-sphere catalogue: exact settlement of the work
// the device left undecided, canonical order and exact level classes.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "pipeline.hpp"
#include "supports.hpp"
#include "timing.hpp"

namespace mhgp {

void supports_device(Context &c, int k_max, std::vector<EventRec> &events,
                     std::vector<FallbackRec> &fbs, std::vector<uint32_t> &fb_region,
                     mhgp_catalog_info *info);

namespace {

struct ExactJob {
  const ExactCloud &ex;
  int s_max;
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
    int64_t witness = -1;
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
    e.level = S.lnum.to_double() / S.lden.to_double() * std::ldexp(1.0, 2 * ex.e0());
    e.level_err = 1e-12f;
    out.push_back(e);
  }
};

// exact squared level of a record as num/den (scaled integer units)
void exact_level(const ExactCloud &ex, const EventRec &e, BigInt &num, BigInt &den) {
  int64_t ids[4];
  int m = e.meta & 0xff;
  for (int k = 0; k < m; ++k) ids[k] = e.ids[k];
  ExactSupport S;
  S.analyze(ex, ids, m);
  num = S.lnum;
  den = S.lden;
}

}  // namespace

void supports(Context &c, int k_max, mhgp_catalog_info *info) {
  if (k_max < 1 || k_max > 10) throw std::invalid_argument("k_max must lie in 1..10");
  *info = mhgp_catalog_info{};
  std::vector<EventRec> events;
  std::vector<FallbackRec> fbs;
  std::vector<uint32_t> region;
  supports_device(c, k_max, events, fbs, region, info);
  StageTimer tm(c.stream);
  CatalogBuffers &B = *c.catalog;
  const int smax = B.s_max;
  // host exact settlement
  if (!fbs.empty()) {
    const ExactCloud &ex = exact_cloud(c);
    ExactJob job{ex, smax, events};
    for (const FallbackRec &f : fbs) {
      if (f.region_off == kInvalid) throw std::runtime_error("fallback without a published region");
      const uint32_t *reg = region.data() + f.region_off;
      const uint32_t len = f.region_len;
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
  }
  tm.mark("catalogue host fallback");
  // canonical order: exact level, rank, support ids
  std::vector<int64_t> order(events.size());
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    if (events[a].level != events[b].level) return events[a].level < events[b].level;
    return a < b;
  });
  std::vector<int64_t> cls(events.size());
  std::vector<EventRec> sorted;
  sorted.reserve(events.size());
  int64_t next_class = 0;
  const ExactCloud *ex = nullptr;
  for (size_t i = 0; i < order.size();) {
    // group records whose FP64 levels are not provably ordered
    size_t j = i + 1;
    double hi = events[order[i]].level * (1.0 + 2.0 * events[order[i]].level_err);
    while (j < order.size() && events[order[j]].level * (1.0 - 2.0 * events[order[j]].level_err) <= hi) {
      hi = std::max(hi, events[order[j]].level * (1.0 + 2.0 * events[order[j]].level_err));
      j++;
    }
    std::vector<int64_t> grp(order.begin() + i, order.begin() + j);
    std::vector<int64_t> gcls(grp.size(), 0);
    if (grp.size() > 1) {
      if (!ex) ex = &exact_cloud(c);
      std::vector<BigInt> num(grp.size()), den(grp.size());
      for (size_t k = 0; k < grp.size(); ++k) exact_level(*ex, events[grp[k]], num[k], den[k]);
      std::vector<size_t> idx(grp.size());
      std::iota(idx.begin(), idx.end(), 0);
      std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        int cmp = compare_levels(num[a], den[a], num[b], den[b]);
        return cmp != 0 ? cmp < 0 : a < b;
      });
      std::vector<int64_t> g2;
      int64_t cc = 0;
      for (size_t k = 0; k < idx.size(); ++k) {
        if (k > 0 && compare_levels(num[idx[k - 1]], den[idx[k - 1]], num[idx[k]], den[idx[k]]) != 0) cc++;
        g2.push_back(grp[idx[k]]);
        gcls[k] = cc;
      }
      grp = g2;
    }
    // inside one exact level: rank, then support ids
    size_t k0 = 0;
    while (k0 < grp.size()) {
      size_t k1 = k0;
      while (k1 < grp.size() && gcls[k1] == gcls[k0]) k1++;
      std::sort(grp.begin() + k0, grp.begin() + k1, [&](int64_t a, int64_t b) {
        const EventRec &x = events[a], &y = events[b];
        uint32_t rx = (x.meta >> 8) & 0xff, ry = (y.meta >> 8) & 0xff;
        if (rx != ry) return rx < ry;
        for (int q = 0; q < 4; ++q)
          if (x.ids[q] != y.ids[q]) return x.ids[q] < y.ids[q];
        return false;
      });
      for (size_t k = k0; k < k1; ++k) {
        sorted.push_back(events[grp[k]]);
        cls[sorted.size() - 1] = next_class + gcls[k];
      }
      k0 = k1;
    }
    next_class += (grp.empty() ? 0 : gcls.back() + 1);
    i = j;
  }
  tm.mark("catalogue canonical sort");
  B.host_events = std::move(sorted);
  B.level_class = std::move(cls);
  int64_t ev = 0, dg = 0;
  for (const EventRec &e : B.host_events) ((e.meta >> 24) & 1 ? dg : ev)++;
  info->event_count = ev;
  info->diagnostic_count = dg;
  info->s_max = smax;
  info->status = 0;
  B.info = *info;
  B.valid = true;
}

void supports_fetch(Context &c, uint32_t *ids, uint32_t *meta, uint32_t *interior, double *level,
                    int64_t *level_class) {
  if (!c.catalog || !c.catalog->valid) throw StateError("no catalogue has been built");
  const CatalogBuffers &B = *c.catalog;
  const int width = std::max(B.s_max - 1, 1);
  for (size_t i = 0; i < B.host_events.size(); ++i) {
    const EventRec &e = B.host_events[i];
    if (ids)
      for (int k = 0; k < 4; ++k) ids[4 * i + k] = e.ids[k];
    if (meta) meta[i] = e.meta;
    if (interior)
      for (int k = 0; k < width; ++k) interior[width * i + k] = k < kMaxInterior ? e.interior[k] : kInvalid;
    if (level) level[i] = e.level;
    if (level_class) level_class[i] = B.level_class[i];
  }
}

}  // namespace mhgp
