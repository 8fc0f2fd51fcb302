// Continuous-batching request server over the per-cloud pipeline.
//
// Worker slots each own a context (and so a HIP stream); requests that arrive
// while others run are admitted as soon as a slot is free and their
// predicted device memory fits under the cap, so the device work of several
// clouds overlaps.  Admission scans the queue in arrival order and takes the
// first request that fits, so a large cloud does not block small ones behind
// it (it is taken as soon as enough memory is free).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "pipeline.hpp"
#include "supports.hpp"
#include "tower_state.hpp"

namespace mhgp {

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t ticket;
  std::vector<double> xyz;
  int k_max;
  uint32_t region_limit;  // the server's enumeration budget when it was submitted
  int64_t predicted;
  Clock::time_point arrival, start, end;
  mhgp_request_result result{};
  TowerState tower;  // the request's forests, concatenated by fetch
  bool done = false;
};

}  // namespace

// Predicted device bytes of one cloud: catalogue records dominate
// (about 640 per point at K_max = 10, 72 B each) plus the LBVH and gathers.
int64_t predict_bytes(int64_t n, int k_max) {
  const int64_t per_point_records = k_max >= 10 ? 640 : 64 * (k_max + 1);
  return n * (per_point_records * (int64_t)(sizeof(EventRec) + 16) + 1024) + (64ll << 20);
}

struct Server {
  int device = 0;
  int64_t cap = 0;
  uint32_t region_limit = 2048;  // enumeration budget of the requests submitted from now on
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_start;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Request>> all;
  int64_t next_ticket = 1;
  // device bytes reserved per slot: what its context retains between
  // requests, or, while it serves one, the larger of that and the request's
  // prediction (contexts keep their grown buffers for the next request)
  std::vector<int64_t> reserved;
  int started = 0, live = 0, start_failures = 0;
  bool stopping = false;
  std::string error;

  int64_t reserved_total() const {
    int64_t t = 0;
    for (int64_t r : reserved) t += r;
    return t;
  }
  void worker_main(int slot);
  void fail_queue_locked(int code);
};

// A fresh context on the slot's device with its own stream.
static std::unique_ptr<Context> fresh_context(int device) {
  auto c = std::make_unique<Context>();
  c->device = device;
  c->tower_lane_count = 1;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return nullptr;
  return c;
}

// Requests nobody can serve any more end with MHGP_E_DEVICE (caller holds mu).
void Server::fail_queue_locked(int code) {
  const auto now = Clock::now();
  for (auto &r : queue) {
    r->result.code = code;
    r->start = r->end = now;
    r->result.latency_ms = std::chrono::duration<double, std::milli>(now - r->arrival).count();
    r->done = true;
  }
  queue.clear();
}

void Server::worker_main(int slot) {
  std::unique_ptr<Context> ctx = fresh_context(device);
  {
    std::lock_guard<std::mutex> g(mu);
    started++;
    if (ctx) live++;
    else {
      start_failures++;
      error = "worker could not create its stream";
    }
  }
  cv_start.notify_all();
  if (!ctx) return;
  while (true) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> g(mu);
      // admission: the slot's share after taking r must fit beside the other
      // slots' reservations; a request larger than the cap runs alone
      auto fits = [&](const Request &q) {
        const int64_t others = reserved_total() - reserved[slot];
        const int64_t mine = std::max(reserved[slot], q.predicted);
        return others + mine <= cap || others == 0;
      };
      cv_work.wait(g, [&] {
        if (stopping) return true;
        for (auto &q : queue) {
          if (fits(*q)) return true;
          // the retained buffers alone are in the way: they are released below
          if (reserved_total() - reserved[slot] + q->predicted <= cap) return true;
        }
        return false;
      });
      if (stopping && queue.empty()) break;
      bool trim = false;
      for (auto it = queue.begin(); it != queue.end(); ++it) {
        if (fits(**it)) {
          r = *it;
          queue.erase(it);
          break;
        }
        if (reserved_total() - reserved[slot] + (*it)->predicted <= cap) trim = true;
      }
      if (!r && trim) {
        // only what this slot retains from earlier requests is in the way:
        // release it (outside the lock) and scan the queue again
        reserved[slot] = 0;
        g.unlock();
        ctx.reset();
        ctx = fresh_context(device);
        g.lock();
      }
      if (!ctx) {
        live--;
        error = "worker could not recreate its stream";
        if (live == 0) fail_queue_locked(MHGP_E_DEVICE);
        cv_done.notify_all();
        cv_work.notify_all();
        return;
      }
      if (!r) continue;
      reserved[slot] = std::max(reserved[slot], r->predicted);
    }
    r->start = Clock::now();
    int status = 0;
    Context &c = *ctx;
    try {
      MHGP_HIP(hipSetDevice(device));
      c.region_limit = r->region_limit;
      upload_host(c, r->xyz.data(), (int64_t)(r->xyz.size() / 3));
      std::vector<uint32_t> edges(2 * std::max<int64_t>(c.n - 1, 0));
      emst_host(c, edges.data());
      mhgp_catalog_info ci{};
      supports(c, r->k_max, &ci);
      r->result.event_count = ci.event_count;
      if (ci.status != 0) {
        // budget stop in the catalogue: reported as conditional, with no tower
        r->result.status = 1;
      } else {
        mhgp_tower_info ti{};
        tower(c, r->k_max, &ti);
        r->tower = std::move(*c.tower);  // the context's tower is rebuilt by its next request
        c.tower->valid = false;
        r->result.node_count = ti.node_count;
        r->result.child_count = ti.child_count;
        r->result.status = ti.status;
      }
    } catch (const DuplicateError &e) {
      status = MHGP_E_DUPLICATE;
    } catch (const NonFiniteError &e) {
      status = MHGP_E_NONFINITE;
    } catch (const std::invalid_argument &e) {
      status = MHGP_E_ARGUMENT;
    } catch (const UnsupportedError &e) {
      status = MHGP_E_UNSUPPORTED;
    } catch (const CapacityError &e) {
      status = MHGP_E_CAPACITY;
    } catch (const HipError &e) {
      status = MHGP_E_DEVICE;
      c.poisoned = true;
      std::lock_guard<std::mutex> g(mu);
      error = e.what();
    } catch (const std::exception &e) {
      status = MHGP_E_INTERNAL;
      c.poisoned = true;
      std::lock_guard<std::mutex> g(mu);
      error = e.what();
    }
    // fail closed: a context whose device state may be inconsistent never
    // serves again; the slot continues on a fresh one
    const bool retire = c.poisoned;
    const int64_t retained = retire ? 0 : device_bytes(c);
    if (retire) {
      ctx.reset();
      ctx = fresh_context(device);
    }
    r->end = Clock::now();
    {
      std::lock_guard<std::mutex> g(mu);
      reserved[slot] = retained;
      r->result.code = status;
      r->result.queue_ms = std::chrono::duration<double, std::milli>(r->start - r->arrival).count();
      r->result.service_ms = std::chrono::duration<double, std::milli>(r->end - r->start).count();
      r->result.latency_ms = std::chrono::duration<double, std::milli>(r->end - r->arrival).count();
      r->done = true;
      if (!ctx) {
        live--;
        error = "worker could not recreate its stream after a device error";
        if (live == 0) fail_queue_locked(MHGP_E_DEVICE);
      }
    }
    cv_done.notify_all();
    cv_work.notify_all();
    if (!ctx) return;
  }
  (void)hipStreamSynchronize(ctx->stream);
}

}  // namespace mhgp

using mhgp::Server;

extern "C" {

struct mhgp_server {
  Server s;
};

int mhgp_server_create(int device, int slots, int64_t memory_cap_bytes, mhgp_server **out) {
  if (!out || slots < 1 || slots > 64) return MHGP_E_ARGUMENT;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return MHGP_E_DEVICE;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return MHGP_E_DEVICE;
  auto *h = new mhgp_server();
  h->s.device = device;
  h->s.cap = memory_cap_bytes > 0 ? memory_cap_bytes : (int64_t)(0.8 * (double)total_b);
  h->s.reserved.assign(slots, 0);
  for (int i = 0; i < slots; ++i) h->s.workers.emplace_back([h, i] { h->s.worker_main(i); });
  bool failed = false;
  {
    std::unique_lock<std::mutex> g(h->s.mu);
    h->s.cv_start.wait(g, [&] { return h->s.started == slots; });
    failed = h->s.start_failures > 0;
  }
  if (failed) {
    mhgp_server_destroy(h);
    return MHGP_E_DEVICE;
  }
  *out = h;
  return MHGP_OK;
}

int mhgp_server_set_region_limit(mhgp_server *srv, int64_t limit) {
  if (!srv || limit < 640 || limit > 2048) return MHGP_E_ARGUMENT;
  std::lock_guard<std::mutex> g(srv->s.mu);
  srv->s.region_limit = (uint32_t)limit;
  return MHGP_OK;
}

int mhgp_server_submit(mhgp_server *srv, const double *xyz, int64_t n, int k_max, int64_t *ticket) {
  if (!srv || !xyz || n <= 0 || k_max < 1 || k_max > 10 || !ticket) return MHGP_E_ARGUMENT;
  auto r = std::make_shared<mhgp::Request>();
  r->xyz.assign(xyz, xyz + 3 * n);
  r->k_max = k_max;
  r->predicted = mhgp::predict_bytes(n, k_max);
  r->arrival = mhgp::Clock::now();
  {
    std::lock_guard<std::mutex> g(srv->s.mu);
    r->region_limit = srv->s.region_limit;
    if (srv->s.live == 0) return MHGP_E_DEVICE;
    r->ticket = srv->s.next_ticket++;
    srv->s.queue.push_back(r);
    srv->s.all[r->ticket] = r;
    *ticket = r->ticket;
  }
  srv->s.cv_work.notify_all();
  return MHGP_OK;
}

int mhgp_server_wait(mhgp_server *srv, int64_t ticket, mhgp_request_result *result) {
  if (!srv || !result) return MHGP_E_ARGUMENT;
  std::unique_lock<std::mutex> g(srv->s.mu);
  auto it = srv->s.all.find(ticket);
  if (it == srv->s.all.end()) return MHGP_E_ARGUMENT;
  auto r = it->second;
  srv->s.cv_done.wait(g, [&] { return r->done || srv->s.live == 0; });
  if (!r->done) return MHGP_E_DEVICE;  // no live worker is left to serve it
  *result = r->result;
  return MHGP_OK;
}

int mhgp_server_fetch(mhgp_server *srv, int64_t ticket, int32_t *order, int32_t *kind,
                      int64_t *level_class, double *level, int64_t *child_offset,
                      int64_t *children) {
  if (!srv) return MHGP_E_ARGUMENT;
  std::lock_guard<std::mutex> g(srv->s.mu);
  auto it = srv->s.all.find(ticket);
  if (it == srv->s.all.end() || !it->second->done) return MHGP_E_STATE;
  auto &r = *it->second;
  r.tower.copy_out(order, kind, level_class, level, child_offset, children);
  return MHGP_OK;
}

int mhgp_server_release(mhgp_server *srv, int64_t ticket) {
  if (!srv) return MHGP_E_ARGUMENT;
  std::lock_guard<std::mutex> g(srv->s.mu);
  srv->s.all.erase(ticket);
  return MHGP_OK;
}

void mhgp_server_destroy(mhgp_server *srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> g(srv->s.mu);
    srv->s.stopping = true;
  }
  srv->s.cv_work.notify_all();
  for (auto &t : srv->s.workers) t.join();
  delete srv;
}

}  // extern "C"
