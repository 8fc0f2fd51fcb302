// hgp_reduced merge-forest tower from the critical-sphere catalogue.
//
// Direct Morse construction (reference morsehgp3d/src/cpu/hierarchy/
// direct_morse_forest_journal.cpp:1042 and the semantics of the exhaustive
// Gamma reduction, reference/morsehgp3d_oracle/hierarchy.py build_merge_forest):
//  * order k minima: critical spheres of closed rank k (order 1: the points);
//  * order k saddles: critical spheres of closed rank k+1; the strict arms of
//    a saddle are its closed label minus one shell point;
//  * an arm is resolved to the minimum of its component by a descent through
//    facets of strictly decreasing miniball: while the closed miniball of the
//    facet F holds a point x outside F, the facet F + x - y (y a point of the
//    miniball's support, x strictly inside) has a smaller miniball and lies in
//    the same component;
//  * saddles of one exact level are contracted together: a group of carriers
//    with no prior reduced root creates a reduced birth (order k >= 2), one
//    root grows silently, two or more roots create a multifusion.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <exception>
#include <mutex>
#include <thread>
#include <numeric>
#include <unordered_map>

#include "pipeline.hpp"
#include "predicates.hpp"
#include "supports.hpp"
#include "timing.hpp"
#include "tower_gpu.hpp"
#include "tower_state.hpp"

namespace mhgp {

namespace {

struct LabelHash {
  size_t operator()(const std::vector<uint32_t> &v) const {
    uint64_t h = 1469598103934665603ull;
    for (uint32_t x : v) { h ^= x; h *= 1099511628211ull; }
    return (size_t)h;
  }
};

struct Grid {
  double lo[3] = {0, 0, 0};
  double cell = 1;
  int64_t dims[3] = {1, 1, 1};
  std::vector<int64_t> start;
  std::vector<uint32_t> ids;
  void build(const std::vector<double> &xyz, int64_t n) {
    double hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
    for (int64_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], xyz[3 * i + a]); hi[a] = std::max(hi[a], xyz[3 * i + a]); }
    double vol = 1;
    for (int a = 0; a < 3; ++a) vol *= std::max(hi[a] - lo[a], 1e-300);
    cell = std::max(std::cbrt(vol / (double)n) * 1.5, 1e-300);
    for (int a = 0; a < 3; ++a) dims[a] = std::min<int64_t>(std::max<int64_t>(1, (int64_t)((hi[a] - lo[a]) / cell) + 1), 1024);
    for (int a = 0; a < 3; ++a) cell = std::max(cell, (hi[a] - lo[a]) / (double)dims[a] * 1.000001);
    const int64_t cells = dims[0] * dims[1] * dims[2];
    start.assign(cells + 1, 0);
    std::vector<int64_t> key(n);
    for (int64_t i = 0; i < n; ++i) { key[i] = cell_of(&xyz[3 * i]); start[key[i] + 1]++; }
    for (int64_t c = 0; c < cells; ++c) start[c + 1] += start[c];
    ids.resize(n);
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < n; ++i) ids[fill[key[i]]++] = (uint32_t)i;
  }
  int64_t coord(double v, int a) const {
    int64_t c = (int64_t)std::floor((v - lo[a]) / cell);
    return std::min<int64_t>(std::max<int64_t>(c, 0), dims[a] - 1);
  }
  int64_t cell_of(const double *p) const { return (coord(p[0], 0) * dims[1] + coord(p[1], 1)) * dims[2] + coord(p[2], 2); }
  template <class F>
  void ball(const double c[3], double r, F &&f) const {
    int64_t a0[3], a1[3];
    for (int a = 0; a < 3; ++a) { a0[a] = coord(c[a] - r, a); a1[a] = coord(c[a] + r, a); }
    for (int64_t i = a0[0]; i <= a1[0]; ++i)
      for (int64_t j = a0[1]; j <= a1[1]; ++j)
        for (int64_t k = a0[2]; k <= a1[2]; ++k) {
          int64_t cc = (i * dims[1] + j) * dims[2] + k;
          for (int64_t q = start[cc]; q < start[cc + 1]; ++q) f(ids[q]);
        }
  }
};

struct BallFit {
  int m = 0;
  int64_t sup[4];
  double c[3];
  double r2 = 0;
};

class Descender {
 public:
  Descender(const std::vector<double> &xyz, int64_t n, const ExactCloud &ex) : xyz_(xyz), n_(n), ex_(ex) {
    grid_.build(xyz, n);
  }
  int64_t steps = 0;

  V3 rel(int64_t x, int64_t o) const {
    return {xyz_[3 * x] - xyz_[3 * o], xyz_[3 * x + 1] - xyz_[3 * o + 1], xyz_[3 * x + 2] - xyz_[3 * o + 2]};
  }

  // certified side of x against the circumsphere of the support (exact fallback)
  int side(const int64_t *sup, int m, int64_t x) const {
    for (int k = 0; k < m; ++k)
      if (sup[k] == x) return 0;
    int s = 0;
    if (m == 1) return 1;
    V3 p = rel(x, sup[0]);
    if (m == 2) { PairSphere ps; ps.init(rel(sup[1], sup[0])); s = ps.side(p); }
    else if (m == 3) { TriSphere ts; ts.init(rel(sup[1], sup[0]), rel(sup[2], sup[0])); s = ts.side(p); }
    else { TetSphere tt; tt.init(rel(sup[1], sup[0]), rel(sup[2], sup[0]), rel(sup[3], sup[0])); s = tt.side(p); }
    if (s != 0) return s;
    ExactSupport S;
    if (!S.analyze(ex_, sup, m)) throw std::runtime_error("dependent support in a descent");
    return S.side(ex_, x);
  }

  // FP64 circumball of a support (centre in its affine hull); false if degenerate
  bool circumball(const int64_t *sup, int m, BallFit &b) const {
    b.m = m;
    for (int k = 0; k < m; ++k) b.sup[k] = sup[k];
    const double *p0 = &xyz_[3 * sup[0]];
    if (m == 1) { for (int a = 0; a < 3; ++a) b.c[a] = p0[a]; b.r2 = 0; return true; }
    V3 a = rel(sup[1], sup[0]);
    V3 off;
    if (m == 2) {
      off = {0.5 * a.x, 0.5 * a.y, 0.5 * a.z};
    } else if (m == 3) {
      V3 bb = rel(sup[2], sup[0]);
      double aa = dot(a, a), b2 = dot(bb, bb), ab = dot(a, bb);
      double D = aa * b2 - ab * ab;
      if (!(D > 1e-14 * (aa * b2 + ab * ab))) return false;
      double al = b2 * (aa - ab), be = aa * (b2 - ab);
      off = {(al * a.x + be * bb.x) / (2 * D), (al * a.y + be * bb.y) / (2 * D), (al * a.z + be * bb.z) / (2 * D)};
    } else {
      V3 bb = rel(sup[2], sup[0]), d = rel(sup[3], sup[0]);
      V3 bxd = cross(bb, d), dxa = cross(d, a), axb = cross(a, bb);
      double V = dot(a, bxd);
      if (!(std::fabs(V) > 1e-14 * adot(vabs(a), across(bb, d)))) return false;
      double aa = dot(a, a), b2 = dot(bb, bb), dd = dot(d, d);
      off = {(aa * bxd.x + b2 * dxa.x + dd * axb.x) / (2 * V), (aa * bxd.y + b2 * dxa.y + dd * axb.y) / (2 * V),
             (aa * bxd.z + b2 * dxa.z + dd * axb.z) / (2 * V)};
    }
    b.c[0] = p0[0] + off.x; b.c[1] = p0[1] + off.y; b.c[2] = p0[2] + off.z;
    b.r2 = dot(off, off);
    return true;
  }

  // minimum enclosing ball of a facet: smallest circumball of a subset that
  // encloses the facet; near-equal radii are compared exactly
  BallFit miniball(const std::vector<uint32_t> &F) const {
    std::vector<BallFit> cands;
    const int k = (int)F.size();
    int64_t s[4];
    for (int m = 1; m <= std::min(4, k); ++m) {
      std::vector<int> idx(m);
      std::iota(idx.begin(), idx.end(), 0);
      while (true) {
        for (int q = 0; q < m; ++q) s[q] = F[idx[q]];
        BallFit b;
        if (circumball(s, m, b)) cands.push_back(b);
        int q = m - 1;
        while (q >= 0 && idx[q] == k - m + q) q--;
        if (q < 0) break;
        idx[q]++;
        for (int r = q + 1; r < m; ++r) idx[r] = idx[r - 1] + 1;
      }
    }
    std::sort(cands.begin(), cands.end(), [](const BallFit &a, const BallFit &b) { return a.r2 < b.r2; });
    const BallFit *best = nullptr;
    for (const BallFit &b : cands) {
      if (best && b.r2 > best->r2 * (1 + 1e-9) + 1e-300) break;
      bool encloses = true;
      for (uint32_t x : F)
        if (side(b.sup, b.m, x) > 0) { encloses = false; break; }
      if (!encloses) continue;
      if (!best) { best = &b; continue; }
      // exact tie-break between two enclosing balls of near-equal radius
      ExactSupport A, B;
      A.analyze(ex_, best->sup, best->m);
      B.analyze(ex_, b.sup, b.m);
      if (compare_levels(B.lnum, B.lden, A.lnum, A.lden) < 0) best = &b;
    }
    if (!best) throw std::runtime_error("no enclosing ball found for a facet");
    return *best;
  }

  // closed-ball members outside the facet; strictly inside ones first
  void extras(const BallFit &b, const std::vector<uint32_t> &F, std::vector<uint32_t> &inside,
              std::vector<uint32_t> &on) const {
    inside.clear();
    on.clear();
    const double r = std::sqrt(b.r2) * (1 + 1e-9) + 1e-300;
    grid_.ball(b.c, r, [&](uint32_t x) {
      if (std::binary_search(F.begin(), F.end(), x)) return;
      double d2 = 0;
      for (int a = 0; a < 3; ++a) { double t = xyz_[3 * x + a] - b.c[a]; d2 += t * t; }
      if (d2 > b.r2 * (1 + 1e-8) + 1e-300) return;
      int sd = side(b.sup, b.m, x);
      if (sd < 0) inside.push_back(x);
      else if (sd == 0) on.push_back(x);
    });
  }

  template <class Map>
  int64_t descend(std::vector<uint32_t> F, const Map &minima) {
    std::vector<uint32_t> in, on;
    for (int it = 0; it < 100000; ++it) {
      auto f = minima.find(F);
      if (f != minima.end()) return f->second;
      BallFit b = miniball(F);
      extras(b, F, in, on);
      if (in.empty()) {
        if (on.empty()) throw std::runtime_error("a facet minimum is missing from the catalogue");
        throw std::runtime_error("degenerate descent: only shell points outside the facet");
      }
      const uint32_t x = *std::min_element(in.begin(), in.end());
      int64_t y = b.sup[0];
      for (int q = 1; q < b.m; ++q) y = std::min(y, b.sup[q]);
      F.erase(std::find(F.begin(), F.end(), (uint32_t)y));
      F.insert(std::upper_bound(F.begin(), F.end(), x), x);
      steps++;
    }
    throw std::runtime_error("a descent did not terminate");
  }

 private:
  const std::vector<double> &xyz_;
  int64_t n_;
  const ExactCloud &ex_;
  Grid grid_;
};


}  // namespace

// Level-ordered contraction of one order: carrier DSU over the minima, each
// class owning at most one reduced root.  Node ids are local to the order.
// Union-find cell of one minimum: parent and the forest node owned by the
// component when the cell is a root (-1: none), in one 8-byte word so that a
// find and the owner lookup after it share the cache line.
struct Cell {
  uint32_t parent;
  int32_t owner;
};

inline uint32_t cell_find(std::vector<Cell> &c, uint32_t x) {
  while (c[x].parent != x) {
    c[x].parent = c[c[x].parent].parent;  // path halving
    x = c[x].parent;
  }
  return x;
}

void contract_order(int k, int64_t n, TowerOrderData &D, Forest &F) {
  const bool timing = std::getenv("MHGP_TIMING") != nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  // carrier union-find; each class may own a reduced root
  const int64_t M = k == 1 ? n : (int64_t)D.minima;
  // the device's arm lists index the arrays below: refuse any that leaves them
  for (uint32_t s = 0; s < D.saddles; ++s)
    if (D.arm_base[s] > D.arm_base[s + 1] || D.arm_base[s + 1] - D.arm_base[s] > 4)
      throw std::runtime_error("malformed saddle arms from the device");
  if (D.saddles && D.arm_base[D.saddles] != D.arms) throw std::runtime_error("malformed saddle arms from the device");
  for (uint32_t a = 0; a < D.arms; ++a)
    if ((int64_t)D.carrier[a] >= M) throw std::runtime_error("an arm's carrier lies outside the order's minima");
  std::vector<Cell> cell(M);
  for (int64_t x = 0; x < M; ++x) cell[x] = {(uint32_t)x, -1};
  F.reserve(k == 1 ? (size_t)n + D.saddles / 8 : D.saddles / 8);
  F.acq_node.assign(M, -1);
  F.acq_cls.assign(M, -1);
  if (k == 1)
    for (int64_t p = 0; p < n; ++p) {
      cell[p].owner = (int32_t)F.add(1, 0, -1, 0.0);
      F.acq_node[p] = cell[p].owner;
      F.rep_elem[p] = (int32_t)p;
    }
  // a class without an owner is a single element (classes only merge in a
  // group, which gives them an owner): record its first owner
  auto acquire = [&](uint32_t root, int64_t node, int64_t cls) {
    if (cell[root].owner < 0) {
      F.acq_node[root] = (int32_t)node;
      F.acq_cls[root] = cls;
    }
  };
  std::unordered_map<int64_t, int64_t> bp;
  std::unordered_map<int64_t, uint32_t> efirst;
  for (uint32_t i = 0; i < D.saddles;) {
    uint32_t j = i;
    while (j < D.saddles && D.cls[j] == D.cls[i]) j++;
    // the walk is latency bound: touch the carriers a few saddles ahead
    if (i + 8 < D.saddles)
      for (uint32_t a = D.arm_base[i + 8]; a < D.arm_base[i + 9]; ++a) __builtin_prefetch(&cell[D.carrier[a]]);
    if (j == i + 1) {
      // single saddle at this level: distinct carrier roots are the group
      uint32_t members[4];
      int nm = 0;
      uint32_t last = kInvalid;
      for (uint32_t a = D.arm_base[i]; a < D.arm_base[i + 1]; ++a) {
        if (D.carrier[a] == last) continue;  // same carrier, same root
        last = D.carrier[a];
        const uint32_t r = cell_find(cell, last);
        bool seen = false;
        for (int q = 0; q < nm; ++q) seen |= members[q] == r;
        if (!seen) members[nm++] = r;
      }
      if (nm > 1) {
        int64_t roots[4];
        int nr = 0;
        for (int q = 0; q < nm; ++q)
          if (cell[members[q]].owner >= 0) roots[nr++] = cell[members[q]].owner;
        std::sort(roots, roots + nr);
        nr = (int)(std::unique(roots, roots + nr) - roots);
        int64_t node;
        if (nr == 0) node = F.add(k, 1, D.cls[i], D.level[i]);
        else if (nr == 1) node = roots[0];
        else node = F.add(k, 2, D.cls[i], D.level[i], roots, nr);
        if (nr != 1) F.rep_elem[node] = (int32_t)members[0];
        for (int q = 0; q < nm; ++q) acquire(members[q], node, D.cls[i]);
        for (int q = 1; q < nm; ++q) cell[members[q]].parent = members[0];
        cell[members[0]].owner = (int32_t)node;
      } else if (cell[members[0]].owner < 0) {
        // all arms in one latent minimum cannot happen in general position
        const int64_t node = F.add(k, 1, D.cls[i], D.level[i]);
        F.rep_elem[node] = (int32_t)members[0];
        acquire(members[0], node, D.cls[i]);
        cell[members[0]].owner = (int32_t)node;
      }
      i = j;
      continue;
    }
    // several saddles at one exact level: contract them together; the groups
    // are emitted in the order of their first saddle
    bp.clear();
    efirst.clear();
    auto bfind = [&](int64_t x) {
      while (bp[x] != x) { bp[x] = bp[bp[x]]; x = bp[x]; }
      return x;
    };
    for (uint32_t e = i; e < j; ++e) {
      int64_t first = -1;
      for (uint32_t a = D.arm_base[e]; a < D.arm_base[e + 1]; ++a) {
        const int64_t r = cell_find(cell, D.carrier[a]);
        if (!bp.count(r)) { bp[r] = r; efirst[r] = e; }
        if (first < 0) first = r;
        else { int64_t x = bfind(first), y = bfind(r); if (x != y) bp[y] = x; }
      }
    }
    std::unordered_map<int64_t, std::pair<uint32_t, std::vector<int64_t>>> groups;
    for (auto &kv : bp) {
      auto &g = groups.try_emplace(bfind(kv.first), UINT32_MAX, std::vector<int64_t>{}).first->second;
      g.first = std::min(g.first, efirst[kv.first]);
      g.second.push_back(kv.first);
    }
    std::vector<std::pair<uint32_t, std::vector<int64_t>>> glist;
    for (auto &kv : groups) {
      std::sort(kv.second.second.begin(), kv.second.second.end());
      glist.push_back(std::move(kv.second));
    }
    std::sort(glist.begin(), glist.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (auto &entry : glist) {
      const std::vector<int64_t> &members = entry.second;
      std::vector<int64_t> roots;
      for (int64_t r : members)
        if (cell[r].owner >= 0) roots.push_back(cell[r].owner);
      std::sort(roots.begin(), roots.end());
      roots.erase(std::unique(roots.begin(), roots.end()), roots.end());
      int64_t node;
      if (roots.empty()) node = F.add(k, 1, D.cls[i], D.level[i]);
      else if (roots.size() == 1) node = roots[0];
      else node = F.add(k, 2, D.cls[i], D.level[i], roots);
      if (roots.size() != 1) F.rep_elem[node] = (int32_t)members[0];
      for (int64_t r : members) acquire((uint32_t)r, node, D.cls[i]);
      for (size_t q = 1; q < members.size(); ++q)
        cell[cell_find(cell, (uint32_t)members[q])].parent = cell_find(cell, (uint32_t)members[0]);
      cell[cell_find(cell, (uint32_t)members[0])].owner = (int32_t)node;
    }
    i = j;
  }
  if (timing)
    std::fprintf(stderr, "[mhgp] contract order %2d: %u saddles, %u arms, %u minima, %zu nodes, %.1f ms\n", k,
                 D.saddles, D.arms, D.minima, F.order.size(),
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  F.raw_of_dense = std::move(D.raw_of_dense);
  D = TowerOrderData{};  // release the order's arrays
}

// Vertical map of order k + 1 onto order k (reference: the profile vertical
// map family of reference/morsehgp3d_oracle/hierarchy.py:656-780 and the
// direct vertical journal, direct_morse_vertical_journal.cpp:822).  A node N
// of order k + 1 at level L maps to the root, at the closed cut L, of the
// order-k component that holds the k-subsets of the facets of N's component.
// N's class at L contains a minimum m of order k + 1 (the representative of
// the group that created N); m's closed label is a saddle of order k whose
// first arm (a k-subset of it) descends to an order-k minimum c, touched no
// later than m's level <= L.  So the target is the owner at L of c's class:
// the node that first owned the class of c's contraction element, raised
// through its ancestors while their level is <= L.
static void vertical_order(const Forest &S, const Forest &T, const std::vector<uint32_t> &elem,
                           std::vector<int64_t> &target) {
  const size_t m = S.order.size(), tm = T.order.size();
  target.assign(m, -1);
  auto start = [&](size_t i) -> int64_t {
    const uint32_t e = elem[i];
    if (e == kInvalid || e >= T.acq_node.size() || T.acq_node[e] < 0 || T.acq_cls[e] > S.level_class[i])
      throw std::runtime_error("vertical map: the order below holds no owned class at the node's level");
    return T.acq_node[e];
  };
  // Nodes are created in level order, so node ids ascend with their level
  // classes; a node's parent has a larger id and a level at least its own.
  // Sweep the queries by level: before answering level L, every node of the
  // order below at level <= L absorbs its children's sets, so the set of a
  // node is headed by its highest ancestor at or below L.
  bool sorted_t = true;
  for (size_t x = 1; x < tm && sorted_t; ++x) sorted_t = T.level_class[x] >= T.level_class[x - 1];
  if (sorted_t) {
    std::vector<uint32_t> q(m);
    std::iota(q.begin(), q.end(), 0u);
    bool sorted_s = true;
    for (size_t i = 1; i < m && sorted_s; ++i) sorted_s = S.level_class[i] >= S.level_class[i - 1];
    if (!sorted_s)
      std::stable_sort(q.begin(), q.end(), [&](uint32_t x, uint32_t y) { return S.level_class[x] < S.level_class[y]; });
    std::vector<uint32_t> head(tm);
    std::iota(head.begin(), head.end(), 0u);
    auto find = [&](uint32_t x) {
      while (head[x] != x) {
        head[x] = head[head[x]];
        x = head[x];
      }
      return x;
    };
    size_t p = 0;
    for (uint32_t i : q) {
      const int64_t L = S.level_class[i];
      for (; p < tm && T.level_class[p] <= L; ++p)
        for (int64_t j = T.child_offset[p]; j < T.child_offset[p + 1]; ++j) head[T.children[j]] = (uint32_t)p;
      target[i] = find((uint32_t)start(i));
    }
    return;
  }
  // otherwise: ancestor tables, up[j][x] the 2^j-th ancestor of x (-1: none);
  // levels never decrease towards the root, so the highest ancestor at or
  // below L is found by descending the powers of two
  std::vector<std::vector<int32_t>> up(1, std::vector<int32_t>(tm, -1));
  for (size_t x = 0; x < tm; ++x)
    for (int64_t j = T.child_offset[x]; j < T.child_offset[x + 1]; ++j) up[0][T.children[j]] = (int32_t)x;
  for (;;) {
    const std::vector<int32_t> &prev = up.back();
    std::vector<int32_t> next(tm, -1);
    bool any = false;
    for (size_t x = 0; x < tm; ++x)
      if (prev[x] >= 0) {
        next[x] = prev[prev[x]];
        any |= next[x] >= 0;
      }
    if (!any) break;
    up.push_back(std::move(next));
  }
  for (size_t i = 0; i < m; ++i) {
    const int64_t L = S.level_class[i];
    int64_t x = start(i);
    for (int j = (int)up.size() - 1; j >= 0; --j) {
      const int32_t a = up[j][x];
      if (a >= 0 && T.level_class[a] <= L) x = a;
    }
    target[i] = x;
  }
}

void tower(Context &c, int k_max, mhgp_tower_info *info) {
  if (!c.catalog || !c.catalog->valid) throw StateError("build the catalogue first");
  CatalogBuffers &B = *c.catalog;
  if (k_max < 1 || k_max > B.k_max) throw std::invalid_argument("k_max exceeds the catalogue's k_max");
  // The direct construction needs every relevant sphere to have shell == support;
  // the reference refuses the journal otherwise (decision
  // no_journal_relevant_extra_shell_diagnostics, direct_morse_event_journal.cpp:587-592).
  if (B.info.status != 0)
    throw UnsupportedError("the catalogue is conditional (" + std::to_string(B.info.anchors_open) +
                           " anchors with candidate regions of up to " + std::to_string(B.info.largest_open_region) +
                           " points were not enumerated): the tower needs the complete catalogue");
  if (B.info.diagnostic_count > 0)
    throw UnsupportedError("the catalogue has " + std::to_string(B.info.diagnostic_count) +
                           " relevant extra-shell diagnostics: the direct Morse tower is not defined");
  StageTimer tm(c.stream);
  const int64_t n = c.n;
  const int K = (int)std::min<int64_t>(k_max, n);
  if (!c.tower) c.tower = std::make_unique<TowerState>();
  TowerState &T = *c.tower;
  T = TowerState{};
  std::unique_ptr<Descender> desc;  // host exact descents, built on demand
  std::vector<double> xyz;
  std::vector<TowerOrderData> Ds(K);
  std::vector<Forest> forests(K);
  std::vector<std::exception_ptr> errors(K);
  std::vector<std::thread> workers;
  struct Joiner {
    std::vector<std::thread> &w;
    ~Joiner() {
      for (auto &t : w)
        if (t.joinable()) t.join();
    }
  } joiner{workers};
  tower_rank_index(c);
  TowerWork &W0 = *c.tower_work;
  W0.stream = c.stream;
  const int lanes = std::max(1, std::min(c.tower_lane_count, K));
  while ((int)c.tower_lanes.size() < lanes - 1) {
    auto w = std::make_unique<TowerWork>();
    MHGP_HIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    w->own_stream = true;
    c.tower_lanes.push_back(std::move(w));
  }
  std::mutex mu;  // workers, the host descender
  // Vertical maps of the order pair (o -> o - 1) need both orders contracted:
  // whichever of the two contractions finishes last runs the pair, while the
  // lanes go on with the orders below.  The element gathers share the first
  // lane's scratch (sized here for every order, so that no buffer is freed
  // while the lanes run) and take turns.
  std::vector<char> contracted(K + 2, 0), pair_done(K + 2, 0);
  std::mutex vm_mu, vm_gather_mu;
  {
    uint64_t most = 16;
    for (int o = 2; o <= K; ++o) most = std::max<uint64_t>(most, 2ull * (W0.rank_hi[o] - W0.rank_lo[o]) + 16);
    if (K >= 2) {
      W0.vm_idx.ensure(most);
      W0.vm_out.ensure(most);
    }
  }
  auto vertical_pair = [&](int o) {
    const Forest &S = forests[o - 1];
    const uint32_t base = W0.rank_lo[o], span = W0.rank_hi[o] - base;
    std::vector<uint32_t> pos(S.order.size());
    for (size_t i = 0; i < S.order.size(); ++i) {
      const int32_t d = S.rep_elem[i];
      if (d < 0) throw std::runtime_error("vertical map: a node without a representative element");
      if (!S.raw_of_dense.empty() && (size_t)d >= S.raw_of_dense.size())
        throw std::runtime_error("vertical map: a representative outside the order's elements");
      const uint32_t raw = S.raw_of_dense.empty() ? (uint32_t)d : S.raw_of_dense[d];
      if (raw >= span) throw std::runtime_error("vertical map: a minimum outside the order's rank range");
      pos[i] = base + raw;
    }
    std::vector<uint32_t> elem;
    {
      std::lock_guard<std::mutex> g(vm_gather_mu);
      tower_vm_gather(c, pos, elem);
    }
    vertical_order(forests[o - 1], forests[o - 2], elem, forests[o - 1].vtarget);
  };
  auto after_contract = [&](int k) {
    int ready[2], nready = 0;
    {
      std::lock_guard<std::mutex> g(vm_mu);
      contracted[k] = 1;
      for (int o : {k + 1, k})
        if (o >= 2 && o <= K && contracted[o] && contracted[o - 1] && !pair_done[o]) {
          pair_done[o] = 1;
          ready[nready++] = o;
        }
    }
    for (int q = 0; q < nready; ++q) vertical_pair(ready[q]);
  };
  // The device lanes take the orders from a shared counter, highest first
  // (their contractions are the longest), so one lane's host
  // synchronisations overlap the other's kernels and a lane that finishes
  // early takes the next order; each order's contraction then runs on a
  // host thread of its own.
  std::atomic<int> next_k{K};
  auto lane = [&](TowerWork &W) {
    MHGP_HIP(hipSetDevice(c.device));
    for (int k = next_k.fetch_sub(1); k >= 1; k = next_k.fetch_sub(1)) {
      TowerOrderData &D = Ds[k - 1];
      tower_order_device(c, W, W0, k, D);
      {
        std::lock_guard<std::mutex> g(mu);
        T.descent_steps += D.steps;
      }
      // carriers the device could not settle: exact host descents
      bool any_host = false;
      for (uint32_t a = 0; a < D.arms; ++a) any_host |= D.carrier[a] == kInvalid;
      if (any_host) {
        std::lock_guard<std::mutex> g(mu);
        if (!desc) {
          xyz.resize(3 * n);
          canonical_points(c, xyz.data(), nullptr);
          desc = std::make_unique<Descender>(xyz, n, exact_cloud(c));
        }
        // device labels index the host cloud: refuse any that leaves it
        for (uint32_t x : D.minima_labels)
          if (x >= (uint64_t)n) throw std::runtime_error("malformed minimum label from the device");
        for (uint32_t x : D.saddle_labels)
          if (x != kInvalid && x >= (uint64_t)n) throw std::runtime_error("malformed saddle label from the device");
        std::unordered_map<std::vector<uint32_t>, int64_t, LabelHash> minima;
        for (uint32_t i = 0; i < D.minima; ++i)
          minima.emplace(std::vector<uint32_t>(D.minima_labels.begin() + (size_t)i * k,
                                               D.minima_labels.begin() + (size_t)(i + 1) * k), i);
        for (uint32_t s = 0; s < D.saddles; ++s)
          for (uint32_t a = D.arm_base[s]; a < D.arm_base[s + 1]; ++a) {
            if (D.carrier[a] != kInvalid) continue;
            std::vector<uint32_t> arm;
            const uint32_t removed = D.sup[4 * s + (a - D.arm_base[s])];
            for (int q = 0; q < 11; ++q) {
              uint32_t x = D.saddle_labels[(size_t)s * 11 + q];
              if (x != kInvalid && x != removed) arm.push_back(x);
            }
            std::sort(arm.begin(), arm.end());
            D.carrier[a] = (uint32_t)desc->descend(arm, minima);
          }
        // vertical maps: the saddles' first-arm carriers are the elements
        std::vector<uint32_t> elems(D.saddles);
        for (uint32_t s2 = 0; s2 < D.saddles; ++s2) elems[s2] = D.carrier[D.arm_base[s2]];
        tower_vm_store(c, W.stream, D.vm_first_pos, elems);
      }
      std::lock_guard<std::mutex> g(mu);
      workers.emplace_back([&, k] {
        try {
          contract_order(k, n, Ds[k - 1], forests[k - 1]);
          after_contract(k);
        } catch (...) {
          errors[k - 1] = std::current_exception();
        }
      });
    }
  };
  std::vector<std::exception_ptr> lane_error(lanes);
  std::vector<std::thread> others;
  for (int l = 1; l < lanes; ++l)
    others.emplace_back([&, l] {
      try {
        lane(*c.tower_lanes[l - 1]);
      } catch (...) {
        lane_error[l] = std::current_exception();
      }
    });
  try {
    lane(W0);
  } catch (...) {
    lane_error[0] = std::current_exception();
  }
  for (auto &t : others) t.join();
  for (auto &e : lane_error)
    if (e) {
      for (auto &t : workers)
        if (t.joinable()) t.join();
      std::rethrow_exception(e);
    }
  std::vector<TowerWork *> all{c.tower_work.get()};
  for (auto &w : c.tower_lanes) all.push_back(w.get());
  for (TowerWork *W : all) {
    c.stats.arm_ms += W->arm_ms;
    c.stats.arm_launches += W->arm_launches;
    c.stats.arm_bytes += W->arm_bytes;
    W->arm_ms = 0;
    W->arm_launches = 0;
    W->arm_bytes = 0;
  }
  for (auto &t : workers) t.join();
  for (auto &e : errors)
    if (e) std::rethrow_exception(e);
  tm.mark("tower forests (host)");
  // every pair's vertical map ran beside the lanes
  for (int o = 2; o <= K; ++o)
    if (!pair_done[o]) throw std::logic_error("vertical map of an order pair was not built");
  tm.mark("tower vertical maps (host)");
  // the orders stay separate; fetches concatenate them once
  for (const Forest &F : forests) {
    T.node_count += (int64_t)F.order.size();
    T.child_count += (int64_t)F.children.size();
  }
  T.forests = std::move(forests);
  T.descent_steps += desc ? desc->steps : 0;
  T.orders = K;
  T.level_exp = -2 * c.scale_exp;
  T.status = 0;
  T.valid = true;
  info->node_count = T.node_count;
  info->child_count = T.child_count;
  info->descent_steps = T.descent_steps;
  info->orders = T.orders;
  info->status = T.status;
}

void tower_fetch(Context &c, int32_t *order, int32_t *kind, int64_t *level_class, double *level,
                 int64_t *child_offset, int64_t *children) {
  if (!c.tower || !c.tower->valid) throw StateError("no tower has been built");
  c.tower->copy_out(order, kind, level_class, level, child_offset, children);
}

// Components of one order at a cut: the nodes at or below the level (strictly
// below for a strict cut) whose parent lies above it, as global node ids
// (reference: K1CompactForest::cut_strict / cut_closed,
// morsehgp3d/include/morsehgp3d/hierarchy/k1_forest.hpp:85-91, for every order).
void tower_cut(Context &c, int order, int64_t cls, int closed, int64_t *roots, int64_t *count,
               int64_t *root_of_point) {
  if (!c.tower || !c.tower->valid) throw StateError("no tower has been built");
  const TowerState &T = *c.tower;
  if (order < 1 || order > T.orders) throw std::invalid_argument("order outside the tower");
  if (!count) throw std::invalid_argument("null count");
  if (root_of_point && order != 1) throw std::invalid_argument("point labels exist for order one only");
  int64_t base = 0;
  for (int o = 1; o < order; ++o) base += (int64_t)T.forests[o - 1].order.size();
  const Forest &F = T.forests[order - 1];
  const size_t m = F.order.size();
  std::vector<int64_t> parent(m, -1);
  for (size_t x = 0; x < m; ++x)
    for (int64_t j = F.child_offset[x]; j < F.child_offset[x + 1]; ++j) parent[F.children[j]] = (int64_t)x;
  auto in = [&](int64_t x) { return closed ? F.level_class[x] <= cls : F.level_class[x] < cls; };
  int64_t nr = 0;
  for (size_t x = 0; x < m; ++x)
    if (in((int64_t)x) && (parent[x] < 0 || !in(parent[x]))) {
      if (roots) roots[nr] = base + (int64_t)x;
      nr++;
    }
  *count = nr;
  if (root_of_point) {
    // order one: node p is the birth of point p
    std::vector<int64_t> top(m, -2);
    for (int64_t p = 0; p < c.n; ++p) {
      if (!in(p)) { root_of_point[p] = -1; continue; }
      int64_t x = p;
      std::vector<int64_t> path;
      while (top[x] == -2 && parent[x] >= 0 && in(parent[x])) { path.push_back(x); x = parent[x]; }
      const int64_t r = top[x] >= 0 ? top[x] : x;
      top[x] = r;
      for (int64_t y : path) top[y] = r;
      root_of_point[p] = base + r;
    }
  }
}

void tower_vertical(Context &c, int64_t *target) {
  if (!c.tower || !c.tower->valid) throw StateError("no tower has been built");
  if (!target) throw std::invalid_argument("null output buffer");
  c.tower->copy_vertical(target);
}

}  // namespace mhgp
