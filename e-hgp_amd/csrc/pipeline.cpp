// Host-side orchestration of the device stages.
#include "pipeline.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "stages.hpp"
#include "supports.hpp"
#include "tower_gpu.hpp"
#include "tower_state.hpp"

namespace mhgp {

Context::Context() = default;
Context::~Context() {
  // device buffers first (their owners are members), then the stream
  tower_lanes.clear();
  tower_work.reset();
  tower.reset();
  host.reset();
  emst.reset();
  catalog.reset();
  if (anchor_clock.a) { (void)hipEventDestroy(anchor_clock.a); (void)hipEventDestroy(anchor_clock.b); }
  if (arm_clock.a) { (void)hipEventDestroy(arm_clock.a); (void)hipEventDestroy(arm_clock.b); }
  if (stream) (void)hipStreamDestroy(stream);
}

int64_t device_bytes(const Context &c) {
  int64_t b = c.src.bytes() + c.X.bytes() * 3 + c.source_index.bytes() + c.key0.bytes() * 3 +
              c.perm.bytes() + c.codes.bytes() + c.leaf_id.bytes() + c.leaf_pos.bytes() +
              c.LX.bytes() * 3 + c.path.bytes() + c.nodes.bytes() + c.parent_int.bytes() +
              c.parent_leaf.bytes() + c.wit.bytes() + c.counter.bytes() + c.wide.bytes();
  if (c.catalog) b += catalog_device_bytes(*c.catalog);
  if (c.emst) b += c.emst->bytes();
  if (c.tower_work) b += c.tower_work->bytes();
  for (const auto &w : c.tower_lanes) b += w->bytes();
  return b;
}

void release_device_state(Context &c) {
  MHGP_HIP(hipStreamSynchronize(c.stream));
  c.tower_lanes.clear();
  c.tower_work.reset();
  c.tower.reset();
  c.host.reset();
  c.emst.reset();
  c.catalog.reset();
  for (DevBuf<double> *b : {&c.src, &c.X, &c.Y, &c.Z, &c.bounds, &c.LX, &c.LY, &c.LZ}) b->release();
  for (DevBuf<uint64_t> *b : {&c.key0, &c.key1, &c.key2, &c.codes, &c.path}) b->release();
  for (DevBuf<uint32_t> *b : {&c.source_index, &c.perm, &c.status, &c.bound_ids, &c.leaf_id, &c.leaf_pos,
                              &c.parent_int, &c.parent_leaf, &c.wit, &c.counter})
    b->release();
  c.nodes.release();
  c.wide.release();
  c.n = 0;
  c.lbvh_ready = false;
}

static void after_upload(Context &c, int64_t n) {
  c.n = n;
  c.lbvh_ready = false;
  c.host.reset();
  if (c.catalog) c.catalog->valid = false;
  if (c.tower) c.tower->valid = false;
  if (c.tower_work) {
    // lists of the previous cloud: finished (they read its LBVH) and dropped
    if (c.tower_work->vm_stream) MHGP_HIP(hipStreamSynchronize(c.tower_work->vm_stream));
    c.tower_work->nn_pending = false;
  }
  try {
    canonicalize_and_build(c);
  } catch (const std::invalid_argument &e) {
    c.n = 0;
    std::string m = e.what();
    if (m.find("duplicate") != std::string::npos) throw DuplicateError(m);
    if (m.find("finite") != std::string::npos) throw NonFiniteError(m);
    throw;
  }
}

void upload_host(Context &c, const double *xyz, int64_t n) {
  c.src.ensure(3 * n);
  MHGP_HIP(hipMemcpyAsync(c.src.p, xyz, 3 * n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  after_upload(c, n);
}

void upload_device(Context &c, const double *d_xyz, int64_t n) {
  c.src.ensure(3 * n);
  MHGP_HIP(hipMemcpyAsync(c.src.p, d_xyz, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  after_upload(c, n);
}

void require_cloud(const Context &c) {
  if (!c.lbvh_ready || c.n <= 0) throw StateError("no cloud has been uploaded");
}

void canonical_points(Context &c, double *xyz_out, uint32_t *source_index_out) {
  const int64_t n = c.n;
  std::vector<double> x(n), y(n), z(n);
  MHGP_HIP(hipMemcpyAsync(x.data(), c.X.p, n * 8, hipMemcpyDeviceToHost, c.stream));
  MHGP_HIP(hipMemcpyAsync(y.data(), c.Y.p, n * 8, hipMemcpyDeviceToHost, c.stream));
  MHGP_HIP(hipMemcpyAsync(z.data(), c.Z.p, n * 8, hipMemcpyDeviceToHost, c.stream));
  if (source_index_out)
    MHGP_HIP(hipMemcpyAsync(source_index_out, c.source_index.p, n * 4, hipMemcpyDeviceToHost, c.stream));
  MHGP_HIP(hipStreamSynchronize(c.stream));
  // back to the input unit (an exact power of two)
  const int e = -c.scale_exp;
  if (xyz_out)
    for (int64_t i = 0; i < n; ++i) {
      xyz_out[3 * i] = std::ldexp(x[i], e);
      xyz_out[3 * i + 1] = std::ldexp(y[i], e);
      xyz_out[3 * i + 2] = std::ldexp(z[i], e);
    }
}

static HostMirror &mirror(Context &c) {
  if (!c.host) {
    c.host = std::make_unique<HostMirror>();
    c.host->xyz.resize(3 * c.n);
    canonical_points(c, c.host->xyz.data(), nullptr);
  }
  return *c.host;
}

const ExactCloud &exact_cloud(Context &c) {
  HostMirror &h = mirror(c);
  if (!h.exact_ready) {
    h.exact.load(h.xyz.data(), c.n);
    h.exact_ready = true;
  }
  return h.exact;
}

// Exact k nearest neighbours of one point by a host scan: FP64 distances
// select a candidate window, exact integer distances order it.
static void knn_exact_one(Context &c, int64_t q, int k, uint32_t *out) {
  const HostMirror &h = mirror(c);
  const ExactCloud &ex = exact_cloud(c);
  const int64_t n = c.n;
  std::vector<std::pair<double, int64_t>> d;
  d.reserve(n - 1);
  for (int64_t j = 0; j < n; ++j) {
    if (j == q) continue;
    double s = 0;
    for (int a = 0; a < 3; ++a) {
      double t = h.xyz[3 * j + a] - h.xyz[3 * q + a];
      s += t * t;
    }
    d.emplace_back(s, j);
  }
  std::nth_element(d.begin(), d.begin() + (k - 1), d.end());
  const double cut = d[k - 1].first * (1.0 + 1e-9) + 1e-300;
  std::vector<int64_t> cand;
  for (auto &e : d)
    if (e.first <= cut) cand.push_back(e.second);
  std::sort(cand.begin(), cand.end(), [&](int64_t a, int64_t b) {
    int s = ex.cmp_dist(q, a, b);
    return s != 0 ? s < 0 : a < b;
  });
  for (int i = 0; i < k; ++i) out[i] = (uint32_t)cand[i];
}

void knn_host(Context &c, int k, uint32_t *ids_out) {
  const int64_t n = c.n;
  DevBuf<uint32_t> d_out;
  DevBuf<uint8_t> d_flags;
  d_out.ensure((size_t)n * k);
  d_flags.ensure(n);
  knn_all(c, k, d_out.p, d_flags.p);
  std::vector<uint8_t> flags(n);
  MHGP_HIP(hipMemcpyAsync(ids_out, d_out.p, (size_t)n * k * 4, hipMemcpyDeviceToHost, c.stream));
  MHGP_HIP(hipMemcpyAsync(flags.data(), d_flags.p, n, hipMemcpyDeviceToHost, c.stream));
  MHGP_HIP(hipStreamSynchronize(c.stream));
  for (int64_t q = 0; q < n; ++q)
    if (flags[q]) knn_exact_one(c, q, k, ids_out + q * k);
}

}  // namespace mhgp
