// Exact Euclidean minimum spanning tree by LBVH Borůvka on gfx950.
//
// Reference: morsehgp3d/src/cpu/hierarchy/boruvka.cpp:893 (build_exact_lbvh_boruvka):
// frozen component labels per round, subtrees uniform in the source's own
// component are skipped, and every component takes its lightest outgoing edge
// under kappa = (d^2, min id, max id).  Under kappa the minimum spanning tree
// is unique, so the output equals the complete-graph Kruskal anchor
// (emst.cpp:326).
//
// Device rounds: node tags -> per-point search -> per-component atomic minima
// -> hooking with 2-cycle breaking -> pointer jumping.  FP64 distances that
// are not provably ordered (near ties that are not both exact) are flagged and
// settled on the host with exact integer arithmetic.
#include <algorithm>

#include "pipeline.hpp"
#include "stages.hpp"
#include "timing.hpp"
#include "traverse.hpp"

namespace mhgp {

namespace {

__device__ inline uint64_t edge_key(uint32_t a, uint32_t b) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  return ((uint64_t)lo << 32) | hi;
}

__global__ void k_init_labels(int64_t n, uint32_t *label) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) label[i] = (uint32_t)i;
}

__global__ void k_zero(uint32_t *a, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// Bottom-up: node_label = common label of all leaves below, or kInvalid.
__global__ void k_node_labels(int64_t n, const Node *__restrict__ nodes,
                              const uint32_t *__restrict__ parent_int,
                              const uint32_t *__restrict__ parent_leaf,
                              const uint32_t *__restrict__ leaf_id,
                              const uint32_t *__restrict__ label, uint32_t *node_label,
                              uint32_t *counter) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t node = parent_leaf[p];
  while (node != kInvalid) {
    __threadfence();
    if (atomicAdd(&counter[node], 1u) == 0) return;
    __threadfence();
    const Node nd = nodes[node];
    uint32_t l[2];
    uint32_t ch[2] = {nd.left, nd.right};
    for (int c = 0; c < 2; ++c) {
      uint32_t x = ch[c];
      l[c] = (x & kLeafBit) ? label[leaf_id[x & ~kLeafBit]]
                            : ((volatile uint32_t *)node_label)[x];
    }
    ((volatile uint32_t *)node_label)[node] = (l[0] == l[1]) ? l[0] : kInvalid;
    node = parent_int[node];
  }
}

// Lightest external edge of every point under kappa.
__global__ void __launch_bounds__(128) k_search(int64_t n, const Node *__restrict__ nodes,
                                                const double *__restrict__ LX,
                                                const double *__restrict__ LY,
                                                const double *__restrict__ LZ,
                                                const uint32_t *__restrict__ leaf_id,
                                                const uint32_t *__restrict__ label,
                                                const uint32_t *__restrict__ node_label,
                                                uint32_t *best_v, double *best_d,
                                                uint8_t *best_flag) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t u = leaf_id[p];
  const uint32_t lu = label[u];
  const double q[3] = {LX[p], LY[p], LZ[p]};
  double bd = INFINITY;
  uint32_t bv = kInvalid;
  bool bx = true;
  uint8_t flag = 0;
  uint32_t stack[64];
  int sp = 0;
  if (node_label[0] != lu) stack[sp++] = 0;
  while (sp) {
    const Node &N = nodes[stack[--sp]];
    uint32_t ch[2] = {N.left, N.right};
    double cd[2] = {INFINITY, INFINITY};
    for (int c = 0; c < 2; ++c) {
      uint32_t x = ch[c];
      if (x & kLeafBit) {
        uint32_t lp = x & ~kLeafBit;
        uint32_t v = leaf_id[lp];
        if (label[v] == lu) continue;
        bool ex;
        double d = d2_flag(q, LX[lp], LY[lp], LZ[lp], ex);
        if (d > bd * (1.0 + kSlack) + kSlackAbs) continue;
        bool both = ex && bx;
        bool better;
        if (d == bd && both) {
          better = edge_key(u, v) < edge_key(u, bv);
        } else {
          if (!both && bd < INFINITY && fabs(d - bd) <= kSlack * bd + kSlackAbs) flag = 1;  // order not provable
          better = d < bd || (d == bd && edge_key(u, v) < edge_key(u, bv));
        }
        if (better) { bd = d; bv = v; bx = ex; }
      } else if (node_label[x] != lu) {
        cd[c] = box_d2(q, nodes[x].lo, nodes[x].hi);
      }
    }
    int first = cd[0] <= cd[1] ? 0 : 1;
    for (int o = 1; o >= 0; --o) {
      int c = o ? 1 - first : first;
      if (cd[c] == INFINITY || cd[c] > bd * (1.0 + kSlack) + kSlackAbs) continue;
      if (sp < 64) stack[sp++] = ch[c]; else flag = 2;
    }
  }
  best_v[u] = bv;
  best_d[u] = bd;
  best_flag[u] = flag | (bx ? 0 : 4);
}

__global__ void k_comp_reset(int64_t n, uint64_t *comp_d, uint64_t *comp_e, uint32_t *comp_amb) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  comp_d[i] = ~0ull;
  comp_e[i] = ~0ull;
  comp_amb[i] = 0;
}

__global__ void k_comp_min(int64_t n, const uint32_t *__restrict__ label,
                           const uint32_t *__restrict__ best_v, const double *__restrict__ best_d,
                           uint64_t *comp_d) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n || best_v[u] == kInvalid) return;
  atomicMin((unsigned long long *)&comp_d[label[u]], (unsigned long long)dbits(best_d[u]));
}

// Entries within rounding of a component's FP64 minimum decide whether the
// winner is proven: a single entry in the window is the exact minimum; several
// entries are settled on the device only when all are exact and equal (ties
// then break on the edge key).  Otherwise the component goes to the host.
__global__ void k_comp_tie(int64_t n, const uint32_t *__restrict__ label,
                           const uint32_t *__restrict__ best_v, const double *__restrict__ best_d,
                           const uint8_t *__restrict__ best_flag, const uint64_t *__restrict__ comp_d,
                           uint64_t *comp_e, uint32_t *comp_amb, uint32_t *comp_cw,
                           uint32_t *comp_ce) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n || best_v[u] == kInvalid) return;
  const uint32_t l = label[u];
  double m;
  uint64_t mb = comp_d[l];
  memcpy(&m, &mb, 8);
  const double d = best_d[u];
  if (best_flag[u] & 2) atomicOr(&comp_amb[l], 16u);  // truncated search: best is only a bound
  if (d - m > kSlack * m + kSlackAbs) return;
  atomicAdd(&comp_cw[l], 1u);
  if (best_flag[u] & 3) atomicOr(&comp_amb[l], 1u);
  if (dbits(d) == mb) {
    atomicMin((unsigned long long *)&comp_e[l], (unsigned long long)edge_key((uint32_t)u, best_v[u]));
    if (!(best_flag[u] & 4)) atomicAdd(&comp_ce[l], 1u);
  }
}

__global__ void k_comp_decide(int64_t n, const uint32_t *__restrict__ label,
                              const uint32_t *__restrict__ comp_cw,
                              const uint32_t *__restrict__ comp_ce, uint32_t *comp_amb) {
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n || label[l] != (uint32_t)l) return;
  if (comp_cw[l] >= 2 && comp_cw[l] != comp_ce[l]) comp_amb[l] |= 8u;
}

// Every component root l hooks to the component across its winning edge.
__global__ void k_hook(int64_t n, const uint32_t *__restrict__ label,
                       const uint64_t *__restrict__ comp_e, uint32_t *hook) {
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  if (label[l] != (uint32_t)l) { hook[l] = kInvalid; return; }
  uint64_t e = comp_e[l];
  if (e == ~0ull) { hook[l] = (uint32_t)l; return; }
  uint32_t a = (uint32_t)(e >> 32), b = (uint32_t)e;
  uint32_t other = label[a] == (uint32_t)l ? b : a;
  hook[l] = label[other];
}

__global__ void k_emit(int64_t n, const uint32_t *__restrict__ label,
                       const uint64_t *__restrict__ comp_e, uint32_t *hook, uint32_t *edges,
                       uint32_t *counters) {
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n || label[l] != (uint32_t)l) return;
  uint32_t h = hook[l];
  if (h == (uint32_t)l) return;
  bool mutual = hook[h] == (uint32_t)l;
  if (mutual && (uint32_t)l > h) return;  // the partner emits the shared edge
  uint64_t e = comp_e[l];
  uint32_t slot = atomicAdd(&counters[0], 1u);
  edges[2 * slot] = (uint32_t)(e >> 32);
  edges[2 * slot + 1] = (uint32_t)e;
}

__global__ void k_break_cycles(int64_t n, const uint32_t *__restrict__ label, uint32_t *hook) {
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n || label[l] != (uint32_t)l) return;
  uint32_t h = hook[l];
  if (h != (uint32_t)l && hook[h] == (uint32_t)l && (uint32_t)l < h) hook[l] = (uint32_t)l;
}

__global__ void k_jump(int64_t n, const uint32_t *__restrict__ label, uint32_t *hook) {
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n || label[l] != (uint32_t)l) return;
  uint32_t h = hook[l];
  uint32_t hh = hook[h];
  if (hh != h) hook[l] = hh;
}

__global__ void k_relabel(int64_t n, uint32_t *label, const uint32_t *__restrict__ hook,
                          uint32_t *counters) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  uint32_t r = hook[label[u]];
  label[u] = r;
  if (r == (uint32_t)u) atomicAdd(&counters[1], 1u);
}

// Roots must be component minima so that labels stay "least PointId";
// k_relabel keeps the hook root, which is the smallest label on a 2-cycle
// but not the smallest PointId of the merged set.  Recompute minima.
__global__ void k_min_label(int64_t n, const uint32_t *__restrict__ label, uint32_t *minid) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < n) atomicMin(&minid[label[u]], (uint32_t)u);
}

__global__ void k_fill(uint32_t *a, int64_t n, uint32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_apply_min(int64_t n, uint32_t *label, const uint32_t *__restrict__ minid) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < n) label[u] = minid[label[u]];
}

}  // namespace

// Host exact settlement of the lightest external edge of one point.
static uint64_t exact_best_edge(Context &c, const std::vector<uint32_t> &label, int64_t u,
                                const ExactCloud &ex, const std::vector<double> &xyz) {
  const int64_t n = c.n;
  double bd = INFINITY;
  for (int64_t v = 0; v < n; ++v) {
    if (label[v] == label[u]) continue;
    double s = 0;
    for (int a = 0; a < 3; ++a) { double t = xyz[3 * v + a] - xyz[3 * u + a]; s += t * t; }
    bd = std::min(bd, s);
  }
  int64_t best = -1;
  for (int64_t v = 0; v < n; ++v) {
    if (label[v] == label[u]) continue;
    double s = 0;
    for (int a = 0; a < 3; ++a) { double t = xyz[3 * v + a] - xyz[3 * u + a]; s += t * t; }
    if (s > bd * (1.0 + 1e-9)) continue;
    if (best < 0) { best = v; continue; }
    int cmp = ex.cmp_dist(u, v, best);
    uint64_t kv = ((uint64_t)std::min<int64_t>(u, v) << 32) | (uint64_t)std::max<int64_t>(u, v);
    uint64_t kb = ((uint64_t)std::min<int64_t>(u, best) << 32) | (uint64_t)std::max<int64_t>(u, best);
    if (cmp < 0 || (cmp == 0 && kv < kb)) best = v;
  }
  return ((uint64_t)std::min<int64_t>(u, best) << 32) | (uint64_t)std::max<int64_t>(u, best);
}

// Sort keys of the forest's edges: the bits of the FP64 squared length (a
// non-negative double orders like its bit pattern) with the edge index.
__global__ void k_kappa_keys(int64_t m, const uint32_t *__restrict__ edges, const double *__restrict__ X,
                             const double *__restrict__ Y, const double *__restrict__ Z, uint64_t *key,
                             uint32_t *idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t a = edges[2 * i], b = edges[2 * i + 1];
  const double dx = X[a] - X[b], dy = Y[a] - Y[b], dz = Z[a] - Z[b];
  key[i] = (uint64_t)__double_as_longlong(dx * dx + dy * dy + dz * dz);
  idx[i] = (uint32_t)i;
}

void emst_host(Context &c, uint32_t *edges_out) {
  const int64_t n = c.n;
  if (n <= 1) return;
  if (!c.emst) c.emst = std::make_unique<EmstBuffers>();
  EmstBuffers &B = *c.emst;
  hipStream_t s = c.stream;
  const int T = 256;
  const int G = blocks_for(n, T);
  B.label.ensure(n);
  B.next_label.ensure(n);
  B.node_label.ensure(n);
  B.best_v.ensure(n);
  B.best_d.ensure(n);
  B.best_flag.ensure(n);
  B.comp_d.ensure(n);
  B.comp_e.ensure(n);
  B.comp_amb.ensure(n);
  B.comp_cw.ensure(n);
  B.comp_ce.ensure(n);
  B.edges.ensure(2 * n);
  B.counters.ensure(4);
  DevBuf<uint32_t> &hook = B.hook;
  hook.ensure(n);
  hipLaunchKernelGGL(k_init_labels, dim3(G), dim3(T), 0, s, n, B.label.p);
  MHGP_HIP(hipMemsetAsync(B.counters.p, 0, 4 * sizeof(uint32_t), s));
  int64_t components = n;
  for (int round = 0; components > 1; ++round) {
    if (round > 64) throw std::runtime_error("Borůvka did not converge");
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(n - 1, T)), dim3(T), 0, s, c.counter.p, n - 1);
    hipLaunchKernelGGL(k_node_labels, dim3(G), dim3(T), 0, s, n, c.nodes.p, c.parent_int.p,
                       c.parent_leaf.p, c.leaf_id.p, B.label.p, B.node_label.p, c.counter.p);
    hipLaunchKernelGGL(k_search, dim3(blocks_for(n, 128)), dim3(128), 0, s, n, c.nodes.p, c.LX.p,
                       c.LY.p, c.LZ.p, c.leaf_id.p, B.label.p, B.node_label.p, B.best_v.p,
                       B.best_d.p, B.best_flag.p);
    hipLaunchKernelGGL(k_comp_reset, dim3(G), dim3(T), 0, s, n, B.comp_d.p, B.comp_e.p, B.comp_amb.p);
    hipLaunchKernelGGL(k_comp_min, dim3(G), dim3(T), 0, s, n, B.label.p, B.best_v.p, B.best_d.p,
                       B.comp_d.p);
    hipLaunchKernelGGL(k_zero, dim3(G), dim3(T), 0, s, B.comp_cw.p, n);
    hipLaunchKernelGGL(k_zero, dim3(G), dim3(T), 0, s, B.comp_ce.p, n);
    hipLaunchKernelGGL(k_comp_tie, dim3(G), dim3(T), 0, s, n, B.label.p, B.best_v.p, B.best_d.p,
                       B.best_flag.p, B.comp_d.p, B.comp_e.p, B.comp_amb.p, B.comp_cw.p,
                       B.comp_ce.p);
    hipLaunchKernelGGL(k_comp_decide, dim3(G), dim3(T), 0, s, n, B.label.p, B.comp_cw.p,
                       B.comp_ce.p, B.comp_amb.p);
    MHGP_LAUNCH_CHECK();
    // host exact settlement of ambiguous components (rare)
    std::vector<uint32_t> amb(n);
    MHGP_HIP(hipMemcpyAsync(amb.data(), B.comp_amb.p, n * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipStreamSynchronize(s));
    bool any = false;
    for (int64_t l = 0; l < n && !any; ++l) any = amb[l] != 0;
    if (any) {
      std::vector<uint32_t> label(n);
      std::vector<uint64_t> ce(n);
      copy_wait(label.data(), B.label.p, n * 4, hipMemcpyDeviceToHost, s);
      copy_wait(ce.data(), B.comp_e.p, n * 8, hipMemcpyDeviceToHost, s);
      const ExactCloud &ex = exact_cloud(c);
      const std::vector<double> &xyz = c.host->xyz;
      for (int64_t l = 0; l < n; ++l) {
        if (!amb[l]) continue;
        uint64_t best = ~0ull;
        for (int64_t u = 0; u < n; ++u) {
          if (label[u] != (uint32_t)l) continue;
          uint64_t e = exact_best_edge(c, label, u, ex, xyz);
          if (best == ~0ull) { best = e; continue; }
          int cmp = compare(ex.sq_dist(e >> 32, (uint32_t)e), ex.sq_dist(best >> 32, (uint32_t)best));
          if (cmp < 0 || (cmp == 0 && e < best)) best = e;
        }
        ce[l] = best;
      }
      copy_wait(B.comp_e.p, ce.data(), n * 8, hipMemcpyHostToDevice, s);
    }
    hipLaunchKernelGGL(k_hook, dim3(G), dim3(T), 0, s, n, B.label.p, B.comp_e.p, hook.p);
    hipLaunchKernelGGL(k_emit, dim3(G), dim3(T), 0, s, n, B.label.p, B.comp_e.p, hook.p,
                       B.edges.p, B.counters.p);
    hipLaunchKernelGGL(k_break_cycles, dim3(G), dim3(T), 0, s, n, B.label.p, hook.p);
    for (int j = 0; j < 24; ++j)
      hipLaunchKernelGGL(k_jump, dim3(G), dim3(T), 0, s, n, B.label.p, hook.p);
    MHGP_HIP(hipMemsetAsync(B.counters.p + 1, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_relabel, dim3(G), dim3(T), 0, s, n, B.label.p, hook.p, B.counters.p);
    hipLaunchKernelGGL(k_fill, dim3(G), dim3(T), 0, s, B.next_label.p, n, kInvalid);
    hipLaunchKernelGGL(k_min_label, dim3(G), dim3(T), 0, s, n, B.label.p, B.next_label.p);
    hipLaunchKernelGGL(k_apply_min, dim3(G), dim3(T), 0, s, n, B.label.p, B.next_label.p);
    MHGP_LAUNCH_CHECK();
    uint32_t cnt[2];
    MHGP_HIP(hipMemcpyAsync(cnt, B.counters.p, 8, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipStreamSynchronize(s));
    components = cnt[1];
    if (cnt[0] + components != (uint64_t)n) throw std::runtime_error("Borůvka edge count mismatch");
  }
  StageTimer tm(s);
  // kappa order: FP64 squared lengths sorted on the device; runs of values
  // within rounding of their neighbours are ordered exactly on the host
  const int64_t m = n - 1;
  B.kappa_key.ensure(m);
  B.kappa_idx.ensure(m);
  hipLaunchKernelGGL(k_kappa_keys, dim3(blocks_for(m, T)), dim3(T), 0, s, m, B.edges.p, c.X.p, c.Y.p, c.Z.p,
                     B.kappa_key.p, B.kappa_idx.p);
  MHGP_LAUNCH_CHECK();
  c.sorter.sort(B.kappa_key.p, B.kappa_idx.p, m, 0, 64, s);
  std::vector<uint32_t> e(2 * m), idx(m);
  std::vector<double> len(m);
  MHGP_HIP(hipMemcpyAsync(e.data(), B.edges.p, e.size() * 4, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipMemcpyAsync(idx.data(), B.kappa_idx.p, m * 4, hipMemcpyDeviceToHost, s));
  copy_wait(len.data(), B.kappa_key.p, m * 8, hipMemcpyDeviceToHost, s);
  std::vector<std::pair<double, uint64_t>> ord(m);
  for (int64_t i = 0; i < m; ++i) {
    const uint32_t x = e[2 * (size_t)idx[i]], y = e[2 * (size_t)idx[i] + 1];
    ord[i] = {len[i], ((uint64_t)std::min(x, y) << 32) | std::max(x, y)};
  }
  // a run continues while each value is within 1e-12 (relative) of the one
  // before it, so values in different runs differ by far more than the
  // rounding of the FP64 lengths and their FP64 order is the exact one
  for (int64_t i = 0; i + 1 < m; ++i) {
    int64_t j = i;
    while (j + 1 < m && ord[j + 1].first - ord[j].first <= 1e-12 * ord[j + 1].first) j++;
    if (j > i) {
      const ExactCloud &ex = exact_cloud(c);
      std::sort(ord.begin() + i, ord.begin() + j + 1, [&](const auto &x, const auto &y) {
        int cmp = compare(ex.sq_dist(x.second >> 32, (uint32_t)x.second),
                          ex.sq_dist(y.second >> 32, (uint32_t)y.second));
        return cmp != 0 ? cmp < 0 : x.second < y.second;
      });
      i = j;
    }
  }
  tm.mark("emst kappa sort");
  for (int64_t i = 0; i < m; ++i) {
    edges_out[2 * i] = (uint32_t)(ord[i].second >> 32);
    edges_out[2 * i + 1] = (uint32_t)ord[i].second;
  }
}

}  // namespace mhgp
