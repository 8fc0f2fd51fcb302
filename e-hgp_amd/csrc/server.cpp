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
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Request>> all;
  int64_t next_ticket = 1;
  int64_t in_flight_bytes = 0;
  bool stopping = false;
  std::string error;

  void worker_main();
};

void Server::worker_main() {
  Context c;
  c.device = device;
  c.tower_lane_count = 1;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
    std::lock_guard<std::mutex> g(mu);
    error = "worker could not create its stream";
    return;
  }
  while (true) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> g(mu);
      cv_work.wait(g, [&] {
        if (stopping) return true;
        for (auto &q : queue)
          if (in_flight_bytes + q->predicted <= cap || in_flight_bytes == 0) return true;
        return false;
      });
      if (stopping && queue.empty()) break;
      for (auto it = queue.begin(); it != queue.end(); ++it)
        if (in_flight_bytes + (*it)->predicted <= cap || in_flight_bytes == 0) {
          r = *it;
          queue.erase(it);
          break;
        }
      if (!r) continue;
      in_flight_bytes += r->predicted;
    }
    r->start = Clock::now();
    int status = 0;
    try {
      MHGP_HIP(hipSetDevice(device));
      c.poisoned = false;
      upload_host(c, r->xyz.data(), (int64_t)(r->xyz.size() / 3));
      std::vector<uint32_t> edges(2 * std::max<int64_t>(c.n - 1, 0));
      emst_host(c, edges.data());
      mhgp_catalog_info ci{};
      supports(c, r->k_max, &ci);
      mhgp_tower_info ti{};
      tower(c, r->k_max, &ti);
      r->tower = std::move(*c.tower);  // the context's tower is rebuilt by its next request
      c.tower->valid = false;
      r->result.event_count = ci.event_count;
      r->result.node_count = ti.node_count;
      r->result.child_count = ti.child_count;
      r->result.status = ci.status | ti.status;
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
      std::lock_guard<std::mutex> g(mu);
      error = e.what();
    } catch (const std::exception &e) {
      status = MHGP_E_INTERNAL;
      std::lock_guard<std::mutex> g(mu);
      error = e.what();
    }
    r->end = Clock::now();
    {
      std::lock_guard<std::mutex> g(mu);
      in_flight_bytes -= r->predicted;
      r->result.code = status;
      r->result.queue_ms = std::chrono::duration<double, std::milli>(r->start - r->arrival).count();
      r->result.service_ms = std::chrono::duration<double, std::milli>(r->end - r->start).count();
      r->result.latency_ms = std::chrono::duration<double, std::milli>(r->end - r->arrival).count();
      r->done = true;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
  (void)hipStreamSynchronize(c.stream);
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
  for (int i = 0; i < slots; ++i) h->s.workers.emplace_back([h] { h->s.worker_main(); });
  *out = h;
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
  srv->s.cv_done.wait(g, [&] { return r->done; });
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
