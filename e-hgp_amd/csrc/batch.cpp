// Batched processing of many clouds: groups of clouds run side by side, one
// host thread and one context (stream) each, and the first anchor pass of
// every cloud of a group is one persistent launch over all of them
// (k_anchor_persist, supports.hip).  A group holds as many clouds as fit the
// memory cap by their predicted device bytes (at most `group`).  Each cloud
// then runs what the plain path runs: exact EMST, catalogue (the remaining
// anchor passes per cloud), and the tower unless the catalogue is
// conditional.  Results and towers are kept per cloud.
#include <algorithm>
#include <chrono>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/mhgp.h"
#include "batch.hpp"
#include "pipeline.hpp"
#include "tower_state.hpp"

namespace mhgp {

int64_t predict_bytes(int64_t n, int k_max);  // server.cpp

namespace {

using Clock = std::chrono::steady_clock;

std::unique_ptr<Context> batch_context(int device) {
  auto c = std::make_unique<Context>();
  c->device = device;
  c->tower_lane_count = 1;
  MHGP_HIP(hipSetDevice(device));
  MHGP_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  return c;
}

}  // namespace

struct BatchRun {
  std::vector<mhgp_request_result> results;
  std::vector<TowerState> towers;
  mhgp_batch_info info{};
};

void batch_run(int device, int64_t count, const double *const *xyz, const int64_t *n, const int32_t *k_max,
               int group, int64_t cap, BatchRun &out) {
  if (count < 0 || group < 1 || group > 256) throw std::invalid_argument("invalid batch shape");
  for (int64_t i = 0; i < count; ++i) {
    if (!xyz[i] || n[i] <= 0 || n[i] >= (int64_t)0x7fffffff) throw std::invalid_argument("invalid point array");
    if (k_max[i] < 1 || k_max[i] > 10) throw std::invalid_argument("k_max must lie in 1..10");
  }
  MHGP_HIP(hipSetDevice(device));
  if (cap <= 0) {
    size_t free_b = 0, total_b = 0;
    MHGP_HIP(hipMemGetInfo(&free_b, &total_b));
    cap = (int64_t)(0.8 * (double)total_b);
  }
  out.results.assign(count, mhgp_request_result{});
  out.towers.assign(count, TowerState{});
  out.info = mhgp_batch_info{};
  out.info.clouds = count;
  const auto t_start = Clock::now();
  std::vector<std::unique_ptr<Context>> pool;
  for (int64_t i = 0; i < count;) {
    // a group: consecutive clouds whose predicted bytes fit the cap together
    int64_t j = i, bytes = 0;
    while (j < count && j - i < group) {
      const int64_t b = predict_bytes(n[j], k_max[j]);
      if (j > i && bytes + b > cap) break;
      bytes += b;
      ++j;
    }
    const int G = (int)(j - i);
    while ((int)pool.size() < G) pool.push_back(batch_context(device));
    auto rv = anchor_batch_create(device, G);
    std::vector<std::thread> th;
    for (int t = 0; t < G; ++t)
      th.emplace_back([&, t] {
        const int64_t cloud = i + t;
        mhgp_request_result &r = out.results[cloud];
        const auto a = Clock::now();
        int code = MHGP_OK;
        std::unique_ptr<Context> &cp = pool[t];
        try {
          if (!cp || cp->poisoned) cp = batch_context(device);
          Context &c = *cp;
          MHGP_HIP(hipSetDevice(device));
          c.batch = rv.get();
          c.batch_slot = t;
          upload_host(c, xyz[cloud], n[cloud]);
          std::vector<uint32_t> edges(2 * std::max<int64_t>(c.n - 1, 0));
          emst_host(c, edges.data());
          mhgp_catalog_info ci{};
          try {
            supports(c, k_max[cloud], &ci);
          } catch (...) {
            anchor_batch_leave(*rv, t);
            throw;
          }
          anchor_batch_leave(*rv, t);  // a cloud that had no anchor pass
          c.batch = nullptr;
          r.event_count = ci.event_count;
          if (ci.status != 0) {
            r.status = 1;  // budget stop: conditional, no tower
          } else {
            mhgp_tower_info ti{};
            tower(c, k_max[cloud], &ti);
            out.towers[cloud] = std::move(*c.tower);
            c.tower->valid = false;
            r.node_count = ti.node_count;
            r.child_count = ti.child_count;
            r.status = ti.status;
          }
        } catch (const DuplicateError &) {
          code = MHGP_E_DUPLICATE;
        } catch (const NonFiniteError &) {
          code = MHGP_E_NONFINITE;
        } catch (const std::invalid_argument &) {
          code = MHGP_E_ARGUMENT;
        } catch (const UnsupportedError &) {
          code = MHGP_E_UNSUPPORTED;
        } catch (const CapacityError &) {
          code = MHGP_E_CAPACITY;
        } catch (const HipError &) {
          code = MHGP_E_DEVICE;
          if (cp) cp->poisoned = true;  // fail closed: the slot gets a fresh context
        } catch (const std::exception &) {
          code = MHGP_E_INTERNAL;
          if (cp) cp->poisoned = true;
        }
        anchor_batch_leave(*rv, t);
        if (cp) cp->batch = nullptr;
        r.code = code;
        r.service_ms = std::chrono::duration<double, std::milli>(Clock::now() - a).count();
        r.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - t_start).count();
      });
    for (auto &x : th) x.join();
    const AnchorBatchStats st = anchor_batch_stats(*rv);
    out.info.groups++;
    out.info.anchors += st.anchors;
    out.info.anchor_ms += st.ms;
    i = j;
  }
  for (auto &c : pool)
    if (c && c->stream) (void)hipStreamSynchronize(c->stream);
  out.info.wall_ms = std::chrono::duration<double, std::milli>(Clock::now() - t_start).count();
}

}  // namespace mhgp

extern "C" {

struct mhgp_batch {
  mhgp::BatchRun run;
};

static constexpr int kBatchPathEnabled = 0;

int mhgp_batch_run(int device, int64_t count, const double *const *xyz, const int64_t *n, const int32_t *k_max,
                   int32_t group, int64_t memory_cap_bytes, mhgp_batch **out) {
  if (!out || count < 0 || (count > 0 && (!xyz || !n || !k_max))) return MHGP_E_ARGUMENT;
  *out = nullptr;
  // Switched off: a process running this path ended with a host segmentation
  // fault whose cause is not known yet.  Callers batch through the server
  // (mhgp_server_*) meanwhile.
  if (kBatchPathEnabled == 0) return MHGP_E_UNSUPPORTED;
  auto *b = new mhgp_batch();
  try {
    mhgp::batch_run(device, count, xyz, n, k_max, group > 0 ? group : 32, memory_cap_bytes, b->run);
  } catch (const std::invalid_argument &) {
    delete b;
    return MHGP_E_ARGUMENT;
  } catch (const std::exception &) {
    delete b;
    return MHGP_E_DEVICE;
  }
  *out = b;
  return MHGP_OK;
}

int mhgp_batch_info_get(const mhgp_batch *b, mhgp_batch_info *info) {
  if (!b || !info) return MHGP_E_ARGUMENT;
  *info = b->run.info;
  return MHGP_OK;
}

int mhgp_batch_result(const mhgp_batch *b, int64_t cloud, mhgp_request_result *result) {
  if (!b || !result || cloud < 0 || cloud >= (int64_t)b->run.results.size()) return MHGP_E_ARGUMENT;
  *result = b->run.results[cloud];
  return MHGP_OK;
}

int mhgp_batch_fetch(const mhgp_batch *b, int64_t cloud, int32_t *order, int32_t *kind, int64_t *level_class,
                     double *level, int64_t *child_offset, int64_t *children) {
  if (!b || cloud < 0 || cloud >= (int64_t)b->run.results.size()) return MHGP_E_ARGUMENT;
  const mhgp::TowerState &T = b->run.towers[cloud];
  if (!T.valid) return MHGP_E_STATE;
  T.copy_out(order, kind, level_class, level, child_offset, children);
  return MHGP_OK;
}

void mhgp_batch_destroy(mhgp_batch *b) { delete b; }

}  // extern "C"
