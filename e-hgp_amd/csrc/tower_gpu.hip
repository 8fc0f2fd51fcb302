// Device arm resolution for the hgp_reduced tower.
//
// For order k the minima are the catalogue records of closed rank k and the
// saddles those of closed rank k+1.  Every strict arm (closed label minus one
// shell point) is resolved to the minimum of its component:
//  * the labels of the minima live in an open-addressing hash table;
//  * an arm that is not a minimum descends: Welzl's algorithm gives the FP64
//    miniball of the facet, its support and containment are certified with
//    the predicates of predicates.hpp, an LBVH ball query lists the points of
//    the closed ball, and the facet trades its least support point for the
//    least point strictly inside (the miniball strictly shrinks);
//  * anything the certified filters cannot decide is flagged for the host's
//    exact descent.
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "pipeline.hpp"
#include "predicates.hpp"
#include "scan.hpp"
#include "supports.hpp"
#include "traverse.hpp"
#include "tower_gpu.hpp"
#include "wave.hpp"

namespace mhgp {

namespace {

constexpr int kMaxLabel = 11;

__device__ inline uint64_t label_hash(const uint32_t *ids, int k) {
  uint64_t h = 0x9e3779b97f4a7c15ull ^ (uint64_t)k;
  for (int i = 0; i < k; ++i) {
    h ^= ids[i] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
  }
  return h | 1ull;  // 0 marks an empty slot
}

__device__ inline int event_label(const EventRec &e, uint32_t *out) {
  const int m = e.meta & 0xff, ni = (e.meta >> 16) & 0xff;
  int k = 0;
  for (int i = 0; i < m; ++i) out[k++] = e.ids[i];
  for (int i = 0; i < ni; ++i) out[k++] = e.interior[i];
  for (int i = 1; i < k; ++i) {  // insertion sort
    uint32_t v = out[i];
    int j = i - 1;
    while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; j--; }
    out[j + 1] = v;
  }
  return k;
}

__device__ inline int rank_of(const EventRec &e) { return (e.meta >> 8) & 0xff; }

// Records grouped by closed rank, canonical order kept inside each rank:
// key = rank (0xff for a degenerate record), value = sorted position.
__global__ void k_rank_keys(const EventRec *__restrict__ ev, const uint32_t *__restrict__ order, int64_t n,
                            uint64_t *keys, uint32_t *vals) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const EventRec &e = ev[order[i]];
  keys[i] = ((e.meta >> 24) & 1) ? 0xffull : (uint64_t)rank_of(e);
  vals[i] = (uint32_t)i;
}

__global__ void k_rank_bounds(const uint64_t *__restrict__ keys, int64_t n, uint32_t *lo, uint32_t *hi) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = keys[i];
  if (i == 0 || keys[i - 1] != key) lo[key] = (uint32_t)i;
  if (i == n - 1 || keys[i + 1] != key) hi[key] = (uint32_t)(i + 1);
}

__global__ void k_insert_minima(const EventRec *__restrict__ ev, const uint32_t *__restrict__ order,
                                const uint32_t *__restrict__ pos, int64_t count, int k,
                                uint32_t *labels, uint64_t *table_keys, uint32_t *table_vals,
                                uint64_t mask) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t lab[kMaxLabel];
  int kk = event_label(ev[order[pos[i]]], lab);
  for (int q = 0; q < k; ++q) labels[(size_t)i * k + q] = q < kk ? lab[q] : kInvalid;
  uint64_t h = label_hash(lab, k);
  uint64_t slot = h & mask;
  while (true) {
    unsigned long long prev = atomicCAS((unsigned long long *)&table_keys[slot], 0ull, (unsigned long long)h);
    if (prev == 0ull) { table_vals[slot] = (uint32_t)i; return; }
    slot = (slot + 1) & mask;
  }
}

__device__ int64_t lookup(const uint32_t *F, int k, const uint64_t *table_keys, const uint32_t *table_vals,
                          uint64_t mask, const uint32_t *labels) {
  uint64_t h = label_hash(F, k);
  uint64_t slot = h & mask;
  while (true) {
    uint64_t key = table_keys[slot];
    if (key == 0) return -1;
    if (key == h) {
      uint32_t v = ((volatile uint32_t *)table_vals)[slot];
      bool same = true;
      for (int q = 0; q < k; ++q) same &= labels[(size_t)v * k + q] == F[q];
      if (same) return v;
    }
    slot = (slot + 1) & mask;
  }
}

struct Ball {
  V3 c;
  double r2;
  int m;
  uint32_t sup[4];  // indices into the facet
};

__device__ inline V3 P3(const double *X, const double *Y, const double *Z, uint32_t id) { return {X[id], Y[id], Z[id]}; }

// circumball of 0..4 facet points with centre in their affine hull (FP64);
// point i of the facet is pts[i * stride]
__device__ bool ball_from(const V3 *pts, int stride, const uint32_t *R, int r, Ball &b) {
  b.m = r;
  for (int i = 0; i < r; ++i) b.sup[i] = R[i];
  if (r == 0) { b.c = {0, 0, 0}; b.r2 = -1; return true; }
  const V3 p0 = pts[R[0] * stride];
  if (r == 1) { b.c = p0; b.r2 = 0; return true; }
  const V3 a = sub(pts[R[1] * stride], p0);
  V3 off;
  if (r == 2) {
    off = {0.5 * a.x, 0.5 * a.y, 0.5 * a.z};
  } else if (r == 3) {
    const V3 bb = sub(pts[R[2] * stride], p0);
    double aa = dot(a, a), b2 = dot(bb, bb), ab = dot(a, bb);
    double D = aa * b2 - ab * ab;
    if (!(D > 1e-14 * (aa * b2 + ab * ab))) return false;
    double al = b2 * (aa - ab), be = aa * (b2 - ab);
    off = {(al * a.x + be * bb.x) / (2 * D), (al * a.y + be * bb.y) / (2 * D), (al * a.z + be * bb.z) / (2 * D)};
  } else {
    const V3 bb = sub(pts[R[2] * stride], p0), d = sub(pts[R[3] * stride], p0);
    V3 bxd = cross(bb, d), dxa = cross(d, a), axb = cross(a, bb);
    double V = dot(a, bxd);
    if (!(fabs(V) > 1e-14 * adot(vabs(a), across(bb, d)))) return false;
    double aa = dot(a, a), b2 = dot(bb, bb), dd = dot(d, d);
    off = {(aa * bxd.x + b2 * dxa.x + dd * axb.x) / (2 * V), (aa * bxd.y + b2 * dxa.y + dd * axb.y) / (2 * V),
           (aa * bxd.z + b2 * dxa.z + dd * axb.z) / (2 * V)};
  }
  b.c = {p0.x + off.x, p0.y + off.y, p0.z + off.z};
  b.r2 = dot(off, off);
  return true;
}

__device__ inline bool in_ball(const Ball &b, const V3 &p) {
  if (b.r2 < 0) return false;
  V3 d = sub(p, b.c);
  return dot(d, d) <= b.r2 * (1.0 + 1e-12) + 1e-300;
}

// Welzl's miniball in its fixed-dimension loop form: level r of the loop nest
// is the recursion level with r boundary points (at most 4 in 3-D), visited in
// the same order as the recursive algorithm.  No recursion, so the kernel's
// private stack has a static size.
__device__ bool welzl(const V3 *pts, int stride, int n, Ball &out) {
  uint32_t R[4];
  Ball b;
  ball_from(pts, stride, R, 0, b);
  for (int i = 0; i < n; ++i) {
    if (in_ball(b, pts[i * stride])) continue;
    R[0] = (uint32_t)i;
    if (!ball_from(pts, stride, R, 1, b)) return false;
    for (int j = 0; j < i; ++j) {
      if (in_ball(b, pts[j * stride])) continue;
      R[1] = (uint32_t)j;
      if (!ball_from(pts, stride, R, 2, b)) return false;
      for (int l = 0; l < j; ++l) {
        if (in_ball(b, pts[l * stride])) continue;
        R[2] = (uint32_t)l;
        if (!ball_from(pts, stride, R, 3, b)) return false;
        for (int q = 0; q < l; ++q) {
          if (in_ball(b, pts[q * stride])) continue;
          R[3] = (uint32_t)q;
          if (!ball_from(pts, stride, R, 4, b)) return false;
        }
      }
    }
  }
  out = b;
  return true;
}

// Sphere of a miniball support (1..4 canonical ids), built once per descent
// step; side() is the certified position of a point (0: on a support point
// or undecided, which sends the arm to the host).
struct SupSphere {
  int m;
  uint32_t sid[4];
  V3 o;
  union {
    PairSphere ps;
    TriSphere ts;
    TetSphere tt;
  };
  // certified well-centring: +1 yes, -1 no, 0 undecided
  __device__ int init(const double *X, const double *Y, const double *Z, const uint32_t *ids, int mm) {
    m = mm;
    for (int i = 0; i < 4; ++i) sid[i] = i < m ? ids[i] : kInvalid;
    o = P3(X, Y, Z, sid[0]);
    if (m == 2) { ps.init(sub(P3(X, Y, Z, sid[1]), o)); return 1; }
    if (m == 3) return ts.init(sub(P3(X, Y, Z, sid[1]), o), sub(P3(X, Y, Z, sid[2]), o));
    if (m == 4)
      return tt.init(sub(P3(X, Y, Z, sid[1]), o), sub(P3(X, Y, Z, sid[2]), o), sub(P3(X, Y, Z, sid[3]), o));
    return 1;
  }
  __device__ int side(const double *X, const double *Y, const double *Z, uint32_t q) const {
    for (int i = 0; i < m; ++i)
      if (sid[i] == q) return 0;
    if (m == 1) return 1;
    const V3 p = sub(P3(X, Y, Z, q), o);
    if (m == 2) return ps.side(p);
    if (m == 3) return ts.side(p);
    return tt.side(p);
  }
};

struct ArmParams {
  const EventRec *ev;
  const uint32_t *order;
  const uint32_t *saddle_pos;  // sorted positions of the saddles of this order
  const uint32_t *arm_base;    // first arm of each saddle (exclusive scan of sizes)
  const uint32_t *arm_saddle;  // saddle of each arm
  int64_t arms;
  int k;
  const uint64_t *table_keys;
  const uint32_t *table_vals;
  uint64_t mask;
  const uint32_t *labels;
  const double *X, *Y, *Z;
  const WideNode *wide;
  const double *LX, *LY, *LZ;
  const uint32_t *leaf_id;
  const uint32_t *leaf_pos;
  int64_t n;
  uint32_t *carrier;  // per arm: minimum index, or kInvalid (host)
  unsigned long long *steps;
  // nearest neighbours of every PointId (kNearK ids, nearest first) and the
  // FP64 squared distance to the last of them; nn == nullptr: none
  const uint32_t *nn;
  const double *nn_d2;
};

constexpr int kNearK = 32;

constexpr int kArmBlock = 64;
constexpr int kArmStack = 40;  // LBVH depth is 18 at 50k points, 32 at 200k; overflow goes to the host

// Facet of an arm: the saddle's closed label without the arm's shell point.
// hint (optional) receives the saddle's other shell points, kInvalid padded:
// the likely support of the facet's miniball.
__device__ inline void arm_facet(const ArmParams &A, int64_t arm, uint32_t *F, uint32_t *hint = nullptr) {
  const uint32_t s = A.arm_saddle[arm];
  const int j = (int)(arm - A.arm_base[s]);
  const EventRec &e = A.ev[A.order[A.saddle_pos[s]]];
  uint32_t L[kMaxLabel];
  const int len = event_label(e, L);
  int f = 0;
  for (int i = 0; i < len; ++i)
    if (L[i] != e.ids[j]) F[f++] = L[i];
  if (hint) {
    const int m = e.meta & 0xff;
    int h = 0;
    for (int q = 0; q < m; ++q)
      if (q != j) hint[h++] = e.ids[q];
    while (h < 4) hint[h++] = kInvalid;
  }
}

// Facets to descend, deduplicated: a slot holds (32 hash bits | 1) << 32 |
// arm; an arm whose facet equals a stored arm's takes that arm's carrier
// after the descents.  The probe length is bounded; an arm that finds no slot
// simply descends itself.
struct FacetSet {
  unsigned long long *slots;
  uint64_t mask;
  uint32_t *alias_arm, *alias_rep;
  unsigned long long *alias_count;
};

__device__ inline bool same_facet(const ArmParams &A, int64_t other, const uint32_t *F) {
  uint32_t G[kMaxLabel];
  arm_facet(A, other, G);
  bool same = true;
  for (int q = 0; q < A.k; ++q) same &= G[q] == F[q];
  return same;
}

// Returns the arm already holding this facet, or -1 (this arm descends).
__device__ inline int64_t facet_claim(const ArmParams &A, const FacetSet &D, int64_t arm, const uint32_t *F) {
  const uint64_t h = label_hash(F, A.k);
  const unsigned long long tag = ((h >> 32) | 1ull) << 32;
  const unsigned long long mine = tag | (unsigned long long)(uint32_t)arm;
  uint64_t slot = h & D.mask;
  for (int probe = 0; probe < 64; ++probe) {
    unsigned long long cur = atomicCAS(&D.slots[slot], 0ull, mine);
    if (cur == 0ull) return -1;
    if ((cur & 0xffffffff00000000ull) == tag) {
      const int64_t other = (int64_t)(cur & 0xffffffffull);
      if (same_facet(A, other, F)) return other;
    }
    slot = (slot + 1) & D.mask;
  }
  return -1;
}

// One thread per strict arm: the facet is a minimum (hash hit) or the arm
// is queued for descents (once per distinct facet).  Splitting the two keeps
// the descending lanes of a wave together.  Queued arms carry the Morton
// position of their facet's least point, so the descents can run in spatial
// order.
__global__ void k_arm_lookup(ArmParams A, FacetSet D, uint32_t *queue, uint64_t *qkey,
                             unsigned long long *qcount) {
  const int64_t arm = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool descend = false, aliased = false;
  int64_t rep = -1;
  uint32_t F[kMaxLabel];
  if (arm < A.arms) {
    arm_facet(A, arm, F);
    if (A.k == 1) {  // order one: the arm is a point, its own minimum
      A.carrier[arm] = F[0];
    } else {
      const int64_t hit = lookup(F, A.k, A.table_keys, A.table_vals, A.mask, A.labels);
      A.carrier[arm] = hit >= 0 ? (uint32_t)hit : kInvalid;
      if (hit < 0) {
        rep = facet_claim(A, D, arm, F);
        aliased = rep >= 0;
        descend = !aliased;
      }
    }
  }
  const unsigned long long ai = wave_append(aliased, D.alias_count);
  if (aliased) {
    D.alias_arm[ai] = (uint32_t)arm;
    D.alias_rep[ai] = (uint32_t)rep;
  }
  // wave-aggregated append to the descent queue
  const uint64_t mask = __ballot(descend);
  if (!mask) return;
  const int lane = (int)(threadIdx.x & 63u);
  const int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned long long at = 0;
  if (lane == leader) at = atomicAdd(qcount, (unsigned long long)__popcll(mask));
  at = __shfl(at, leader);
  if (descend) {
    const unsigned long long pos = at + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
    queue[pos] = (uint32_t)arm;
    qkey[pos] = A.leaf_pos[F[0]];
  }
}

__global__ void k_alias_carriers(const uint32_t *__restrict__ alias_arm, const uint32_t *__restrict__ alias_rep,
                                 uint64_t count, uint32_t *carrier) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) carrier[alias_arm[i]] = carrier[alias_rep[i]];
}

// Trade the miniball's least support point for the least point strictly
// inside; false leaves the arm to the host (shell points, or no point inside:
// a minimum missing from the table).
__device__ bool finish_step(const Ball &b, const uint32_t *sid, int k, uint32_t *F, uint32_t *hint, uint32_t xin,
                            bool undecided) {
  if (undecided || xin == kInvalid) return false;
  uint32_t y = sid[0];
  for (int i = 1; i < b.m; ++i) y = min(y, sid[i]);
  // F := F - y + xin, kept sorted
  int w = 0;
  uint32_t G[kMaxLabel];
  bool placed = false;
  for (int i = 0; i < k; ++i) {
    if (F[i] == y) continue;
    if (!placed && xin < F[i]) { G[w++] = xin; placed = true; }
    G[w++] = F[i];
  }
  if (!placed) G[w++] = xin;
  for (int i = 0; i < k; ++i) F[i] = G[i];
  // the next miniball's likely support: this one's, y traded for xin
  for (int h = 0; h < 4; ++h) hint[h] = h < b.m ? (sid[h] == y ? xin : sid[h]) : kInvalid;
  return true;
}

// One descent step of facet F (sorted ids) with the likely support hint:
// F := F - y + x, where x is the least point strictly inside the facet's
// miniball and y its least support point.  false: the certified filters
// cannot decide the step (the host descends the arm).  stk is the lane's
// traversal stack in LDS (stride kArmBlock).
__device__ bool descent_step(const ArmParams &A, uint32_t *F, uint32_t *hint, uint32_t *stk, int lane) {
  const int k = A.k;
  // miniball of the facet; the likely support goes first, so that Welzl's
  // outer loop mostly meets points already inside
  uint32_t ord[kMaxLabel];
  int no = 0;
  for (int h = 0; h < 4; ++h) {
    bool in = hint[h] != kInvalid;
    if (in) {
      in = false;
      for (int i = 0; i < k; ++i) in |= F[i] == hint[h];
    }
    if (in) ord[no++] = hint[h];
  }
  for (int i = 0; i < k; ++i) {
    bool hinted = false;
    for (int h = 0; h < 4; ++h) hinted |= hint[h] == F[i];
    if (!hinted) ord[no++] = F[i];
  }
  V3 pts[kMaxLabel];
  for (int i = 0; i < k; ++i) pts[i] = P3(A.X, A.Y, A.Z, ord[i]);
  Ball b;
  if (!welzl(pts, 1, k, b)) return false;
  uint32_t sid[4];
  for (int i = 0; i < b.m; ++i) sid[i] = ord[b.sup[i]];
  SupSphere S;
  if (S.init(A.X, A.Y, A.Z, sid, b.m) <= 0) return false;
  // every facet point must be certainly inside, or a support point
  bool ok = true;
  for (int i = 0; i < k && ok; ++i) {
    bool on_support = false;
    for (int q = 0; q < b.m; ++q) on_support |= sid[q] == F[i];
    if (!on_support) ok = S.side(A.X, A.Y, A.Z, F[i]) < 0;
  }
  if (!ok) return false;
  // closed-ball points outside the facet
  uint32_t xin = kInvalid;
  bool undecided = false;
  const double q[3] = {b.c.x, b.c.y, b.c.z};
  const double R2 = b.r2 * (1.0 + 1e-9) + 1e-300;
  if (A.nn) {
    // every point of the ball lies within sqrt(R2) + |c - f| of a facet point
    // f; when that radius stays below the distance to f's last listed
    // neighbour, f's neighbour list holds all of them and replaces the
    // traversal (with margins far above the FP64 rounding)
    uint32_t f = F[0];
    double best = INFINITY;
    for (int i = 0; i < k; ++i) {
      const double dx = A.X[F[i]] - q[0], dy = A.Y[F[i]] - q[1], dz = A.Z[F[i]] - q[2];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < best) { best = d2; f = F[i]; }
    }
    const double rho = (sqrt(R2) + sqrt(best)) * (1.0 + 1e-9) + 1e-150;
    if (rho * rho < A.nn_d2[f] * (1.0 - 1e-9)) {
      for (int j = 0; j < kNearK && !undecided; ++j) {
        const uint32_t id = A.nn[(size_t)f * kNearK + j];
        const double dx = A.X[id] - q[0], dy = A.Y[id] - q[1], dz = A.Z[id] - q[2];
        if (dx * dx + dy * dy + dz * dz > R2) continue;
        bool inF = false;
        for (int i = 0; i < k; ++i) inF |= F[i] == id;
        if (inF) continue;
        const int sd = S.side(A.X, A.Y, A.Z, id);
        if (sd < 0) xin = min(xin, id);
        else if (sd == 0) undecided = true;
      }
      return finish_step(b, sid, k, F, hint, xin, undecided);
    }
  }
  // FP32 pruning against outward-rounded child boxes: the radius is
  // inflated by the rounding of the query point and of the FP32 arithmetic
  // (a non-finite bound prunes nothing)
  const float qf[3] = {(float)q[0], (float)q[1], (float)q[2]};
  const double qmax = fmax(fmax(fabs(q[0]), fabs(q[1])), fabs(q[2]));
  const double rb = sqrt(R2) * (1.0 + 1e-6) + qmax * 1e-6 + 1e-30;
  float Tf = (float)(rb * rb * (1.0 + 1e-5));
  if (!(Tf < 3.0e38f)) Tf = INFINITY;
  int sp = 0;
  stk[lane] = 0;
  sp = 1;
  while (sp && !undecided) {
    const WideNode &N = A.wide[stk[(--sp) * kArmBlock + lane]];
    for (int c = 0; c < 4; ++c) {
      float d2 = 0.f;
      for (int a = 0; a < 3; ++a) {
        const float d = fmaxf(fmaxf(N.lo[c][a] - qf[a], qf[a] - N.hi[c][a]), 0.f);
        d2 += d * d;
      }
      const uint32_t x = N.child[c];
      if (d2 > Tf || x == kInvalid) continue;
      if (x & kLeafBit) {
        const uint32_t lp = x & ~kLeafBit;
        const double dx = A.LX[lp] - q[0], dy = A.LY[lp] - q[1], dz = A.LZ[lp] - q[2];
        if (dx * dx + dy * dy + dz * dz > R2) continue;
        const uint32_t id = A.leaf_id[lp];
        bool inF = false;
        for (int i = 0; i < k; ++i) inF |= F[i] == id;
        if (inF) continue;
        const int sd = S.side(A.X, A.Y, A.Z, id);
        if (sd < 0) xin = min(xin, id);
        else if (sd == 0) undecided = true;
      } else if (sp < kArmStack) {
        stk[(sp++) * kArmBlock + lane] = x;
      } else {
        undecided = true;
      }
    }
  }
  return finish_step(b, sid, k, F, hint, xin, undecided);
}

// Descents advance one step per round for every arm still descending, so the
// lanes of a wave do equal work.  Item state lives in SoA arrays of
// capacity cap: facet ids F[q * cap + i], hint H[h * cap + i], origin arm.
struct DescentItems {
  uint32_t *F, *H, *origin;
  uint64_t cap;
};

// Round 0 takes the queued arms (their facets are not minima); later rounds
// first look their facet up among the minima.
__global__ void __launch_bounds__(kArmBlock) k_descend_round(ArmParams A, int round, const uint32_t *queue,
                                                             DescentItems in, uint32_t n_in, DescentItems out,
                                                             unsigned long long *n_out) {
  __shared__ uint32_t stk[kArmStack * kArmBlock];
  const int lane = threadIdx.x;
  const uint64_t i = (uint64_t)blockIdx.x * kArmBlock + lane;
  const int k = A.k;
  bool next = false;
  uint32_t F[kMaxLabel], hint[4], origin = 0;
  if (i < n_in) {
    if (round == 0) {
      origin = queue[i];
      arm_facet(A, origin, F, hint);
    } else {
      origin = in.origin[i];
      for (int q = 0; q < k; ++q) F[q] = in.F[q * in.cap + i];
      for (int h = 0; h < 4; ++h) hint[h] = in.H[h * in.cap + i];
    }
    int64_t hit = -1;
    if (round > 0) hit = lookup(F, k, A.table_keys, A.table_vals, A.mask, A.labels);
    if (hit >= 0) {
      A.carrier[origin] = (uint32_t)hit;
    } else if (round >= 4096 || !descent_step(A, F, hint, stk, lane)) {
      A.carrier[origin] = kInvalid;  // the host descends this arm
    } else {
      next = true;
    }
  }
  const unsigned long long at = wave_append(next, n_out);
  if (next) {
    for (int q = 0; q < k; ++q) out.F[q * out.cap + at] = F[q];
    for (int h = 0; h < 4; ++h) out.H[h * out.cap + at] = hint[h];
    out.origin[at] = origin;
  }
}

__global__ void k_arm_owner(const uint32_t *__restrict__ base, const uint32_t *__restrict__ sizes_incl,
                            int64_t count, uint32_t *arm_saddle) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t b = base[i], e = sizes_incl[i];
  for (uint32_t a = b; a < e; ++a) arm_saddle[a] = (uint32_t)i;
}

__global__ void k_arm_sizes(const EventRec *__restrict__ ev, const uint32_t *__restrict__ order,
                            const uint32_t *__restrict__ pos, int64_t count, uint32_t *sizes) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) sizes[i] = ev[order[pos[i]]].meta & 0xff;
}

__global__ void k_saddle_meta(const EventRec *__restrict__ ev, const uint32_t *__restrict__ order,
                              const uint32_t *__restrict__ cls, const uint32_t *__restrict__ pos,
                              int64_t count, uint32_t *out_cls, double *out_level) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const EventRec &e = ev[order[pos[i]]];
  out_cls[i] = cls[pos[i]];
  out_level[i] = e.level;
}

// Closed labels (support then interior, kInvalid padded to 11) and supports
// of the saddles, for the host's exact descents.
__global__ void k_saddle_labels(const EventRec *__restrict__ ev, const uint32_t *__restrict__ order,
                                const uint32_t *__restrict__ pos, int64_t count, uint32_t *labels,
                                uint32_t *sup) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const EventRec &e = ev[order[pos[i]]];
  const int m = e.meta & 0xff, ni = (e.meta >> 16) & 0xff;
  int t = 0;
  for (int q = 0; q < m; ++q) labels[(size_t)i * kMaxLabel + t++] = e.ids[q];
  for (int q = 0; q < ni; ++q) labels[(size_t)i * kMaxLabel + t++] = e.interior[q];
  while (t < kMaxLabel) labels[(size_t)i * kMaxLabel + t++] = kInvalid;
  for (int q = 0; q < 4; ++q) sup[4 * i + q] = e.ids[q];
}

// Reduction of an order's saddles before the level-ordered contraction.
//
// first[m] is the least saddle with an arm at minimum m.  A carrier m of
// saddle s is fresh when first[m] == s (an untouched one-minimum component),
// owned otherwise (its component already has a reduced root; order one has
// no fresh minima, every point is born at level zero).
//  * Attachment: every fresh carrier joins, from its first saddle on, the
//    component of that saddle's first owned carrier, or of its first fresh
//    carrier when it has none.  The attachment pointers form a forest (a
//    pointer goes to a minimum first touched strictly earlier, or to a root)
//    and its roots are the representatives.
//  * Owned carriers of one saddle on distinct representatives give star
//    edges, keyed by arm index (level order).  A Boruvka spanning forest over
//    them counts, per saddle, the components its owned carriers join beyond
//    the first.
//  * A saddle at an unshared level with owned carriers and no such forest edge
//    changes no root (silent growth or no new connection) and is dropped; the
//    others (births, multifusions, shared levels) go to the host with their
//    carriers replaced by representatives.
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u64(uint64_t *p, int64_t n, uint64_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_identity(uint32_t *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

__global__ void k_first_touch(const uint32_t *__restrict__ arm_saddle, const uint32_t *__restrict__ carrier,
                              int64_t arms, int64_t minima, uint32_t *first, uint32_t *invalid) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= arms) return;
  const uint32_t c = carrier[a];
  if (c == kInvalid || c >= minima) {
    *invalid = 1u;
    return;
  }
  atomicMin(&first[c], arm_saddle[a]);
}

__device__ inline bool carrier_owned(const uint32_t *first, uint32_t c, uint32_t s, int all_owned) {
  return all_owned || first[c] < s;
}

__global__ void k_attach(int64_t S, const uint32_t *__restrict__ base, const uint32_t *__restrict__ incl,
                         const uint32_t *__restrict__ carrier, const uint32_t *__restrict__ first, int all_owned,
                         uint32_t *attach) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const uint32_t b = base[s], e = incl[s];
  uint32_t target = kInvalid, fresh0 = kInvalid;
  for (uint32_t a = b; a < e; ++a) {
    const uint32_t c = carrier[a];
    if (carrier_owned(first, c, (uint32_t)s, all_owned)) { target = c; break; }
    if (fresh0 == kInvalid) fresh0 = c;
  }
  if (target == kInvalid) target = fresh0;
  for (uint32_t a = b; a < e; ++a) {
    const uint32_t c = carrier[a];
    if (c != target && !carrier_owned(first, c, (uint32_t)s, all_owned)) attach[c] = target;
  }
}

// Star edge of arm a: an owned carrier on a representative that neither the
// saddle's first owned carrier nor an earlier owned arm has.
__device__ inline bool arm_edge(const uint32_t *base, const uint32_t *incl, const uint32_t *carrier,
                                const uint32_t *first, const uint32_t *rep, uint32_t s, uint32_t a, int all_owned,
                                uint32_t &r0, uint32_t &r) {
  const uint32_t c = carrier[a];
  if (!carrier_owned(first, c, s, all_owned)) return false;
  r = rep[c];
  uint32_t x = base[s];
  while (!carrier_owned(first, carrier[x], s, all_owned)) ++x;  // a itself is owned
  if (x == a) return false;
  r0 = rep[carrier[x]];
  if (r == r0) return false;
  for (++x; x < a; ++x)
    if (carrier_owned(first, carrier[x], s, all_owned) && rep[carrier[x]] == r) return false;
  (void)incl;
  return true;
}

__global__ void k_edge_flags(const uint32_t *__restrict__ arm_saddle, const uint32_t *__restrict__ base,
                             const uint32_t *__restrict__ incl, const uint32_t *__restrict__ carrier,
                             const uint32_t *__restrict__ first, const uint32_t *__restrict__ rep, int all_owned,
                             int64_t arms, uint32_t *flag) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= arms) return;
  uint32_t r0, r;
  flag[a] = arm_edge(base, incl, carrier, first, rep, arm_saddle[a], (uint32_t)a, all_owned, r0, r) ? 1u : 0u;
}

__global__ void k_edge_list(const uint32_t *__restrict__ arm_saddle, const uint32_t *__restrict__ base,
                            const uint32_t *__restrict__ incl, const uint32_t *__restrict__ carrier,
                            const uint32_t *__restrict__ first, const uint32_t *__restrict__ rep, int all_owned,
                            const uint32_t *__restrict__ scan, int64_t arms, uint32_t *eu, uint32_t *ev,
                            uint32_t *es) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= arms) return;
  const uint32_t prev = a ? scan[a - 1] : 0u;
  if (scan[a] == prev) return;
  const uint32_t s = arm_saddle[a];
  uint32_t r0 = 0, r = 0;
  arm_edge(base, incl, carrier, first, rep, s, (uint32_t)a, all_owned, r0, r);
  eu[prev] = r0;
  ev[prev] = r;
  es[prev] = s;
}

// Boruvka round, part 1: least edge (by index) leaving each component.  The
// edges still joining two components are appended to the next round's list
// (li == nullptr: the list is the full edge list, li[i] = i).
__global__ void k_msf_best(int64_t L, const uint32_t *__restrict__ lu, const uint32_t *__restrict__ lv,
                           const uint32_t *__restrict__ li, const uint32_t *__restrict__ comp, uint32_t *best,
                           uint32_t *ou, uint32_t *ov, uint32_t *oi, uint32_t *count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t u = 0, v = 0, idx = 0, cu = 0, cv = 0;
  if (i < L) {
    u = lu[i];
    v = lv[i];
    idx = li ? li[i] : (uint32_t)i;
    cu = comp[u];
    cv = comp[v];
  }
  const bool live = i < L && cu != cv;
  if (live) {
    // the list runs in increasing index order, so a plain read first lets most
    // edges of a large component skip an atomic that cannot win
    if (idx < best[cu]) atomicMin(&best[cu], idx);
    if (idx < best[cv]) atomicMin(&best[cv], idx);
  }
  // wave-aggregated append
  const uint64_t mask = __ballot(live);
  if (!mask) return;
  const int lane = (int)(threadIdx.x & 63u);
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t at = 0;
  if (lane == leader) at = atomicAdd(count, (uint32_t)__popcll(mask));
  at = __shfl(at, leader);
  if (live) {
    const uint32_t pos = at + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    ou[pos] = u;
    ov[pos] = v;
    oi[pos] = idx;
  }
}

// Part 2: every component root hooks along its least edge, which is a forest
// edge; of a pair choosing the same edge the smaller root stays a root.
// comp is only read here; the hooks are applied by the next kernel.
__global__ void k_msf_hook(int64_t M, const uint32_t *__restrict__ eu, const uint32_t *__restrict__ ev,
                           const uint32_t *__restrict__ comp, const uint32_t *__restrict__ best, uint32_t *hook,
                           uint32_t *msf, uint32_t *changed) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M || comp[m] != (uint32_t)m) return;
  const uint32_t i = best[m];
  if (i == kInvalid) return;
  const uint32_t cu = comp[eu[i]], cv = comp[ev[i]];
  const uint32_t other = cu == (uint32_t)m ? cv : cu;
  msf[i] = 1u;
  if (best[other] == i && (uint32_t)m < other) return;
  hook[m] = other;
  *changed = 1u;
}

__global__ void k_msf_apply(int64_t M, uint32_t *comp, uint32_t *best, uint32_t *hook) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M || comp[m] != (uint32_t)m) return;
  if (hook[m] != kInvalid) comp[m] = hook[m];
  hook[m] = kInvalid;
  best[m] = kInvalid;
}

// Root of m in a forest of parent pointers (a root points to itself, or
// holds kInvalid when nil_root); path halving.  Concurrent threads only ever
// store an ancestor of a node into it, so the forest keeps its roots; the
// roots themselves go to another array.  The walk is bounded by the node
// count: a longer one (a cycle, which a correct forest cannot hold) raises
// *err and returns kInvalid.
__device__ inline uint32_t forest_root(uint32_t *parent, uint32_t m, bool nil_root, int64_t nodes, uint32_t *err) {
  uint32_t x = m;
  for (int64_t step = 0;; ++step) {
    if (step > nodes) {
      *err = 1u;
      return kInvalid;
    }
    const uint32_t p = parent[x];
    if (nil_root ? p == kInvalid : p == x) break;
    const uint32_t g = parent[p];
    if (nil_root ? g == kInvalid : g == p) { x = p; break; }
    parent[x] = g;
    x = g;
  }
  return x;
}

// Roots of the hook forest into the next round's component array.
__global__ void k_msf_flatten(int64_t M, uint32_t *comp, uint32_t *next, uint32_t *err) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  next[m] = forest_root(comp, (uint32_t)m, false, M, err);
}

__global__ void k_msf_count(int64_t E, const uint32_t *__restrict__ msf, const uint32_t *__restrict__ es,
                            uint32_t *mcount) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < E && msf[i]) atomicAdd(&mcount[es[i]], 1u);
}

__global__ void k_reduce_saddles(int64_t S, const uint32_t *__restrict__ base, const uint32_t *__restrict__ incl,
                                 const uint32_t *__restrict__ carrier, const uint32_t *__restrict__ cls,
                                 const uint32_t *__restrict__ first, const uint32_t *__restrict__ mcount,
                                 int all_owned, uint32_t *keep, uint32_t *kept_arms) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const uint32_t b = base[s], e = incl[s];
  const uint32_t c = cls[s];
  const bool shared = (s > 0 && cls[s - 1] == c) || (s + 1 < S && cls[s + 1] == c);
  bool k = true;
  if (!shared && mcount[s] == 0) {
    bool any_owned = false;
    for (uint32_t a = b; a < e; ++a) any_owned |= carrier_owned(first, carrier[a], (uint32_t)s, all_owned);
    k = !any_owned;  // no owned carrier: a birth
  }
  keep[s] = k ? 1u : 0u;
  kept_arms[s] = k ? e - b : 0u;
}

__global__ void k_attach_roots(int64_t M, uint32_t *attach, uint32_t *rep, uint32_t *err) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < M) rep[m] = forest_root(attach, (uint32_t)m, true, M, err);
}

// Minima met by a kept saddle get dense ids (flag, then an inclusive scan).
__global__ void k_flag_used(int64_t S, const uint32_t *__restrict__ keep, const uint32_t *__restrict__ base,
                            const uint32_t *__restrict__ incl, const uint32_t *__restrict__ carrier,
                            const uint32_t *__restrict__ rep, uint32_t *used) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S || !keep[s]) return;
  for (uint32_t a = base[s]; a < incl[s]; ++a) used[rep[carrier[a]]] = 1u;
}

__global__ void k_compact_saddles(int64_t S, const uint32_t *__restrict__ kpos, const uint32_t *__restrict__ kinc,
                                  const uint32_t *__restrict__ base, const uint32_t *__restrict__ incl,
                                  const uint32_t *__restrict__ carrier, const uint32_t *__restrict__ scls,
                                  const double *__restrict__ slev, const uint32_t *__restrict__ rep,
                                  const uint32_t *__restrict__ dense, uint32_t *o_cls, double *o_lev, uint32_t *o_base,
                                  uint32_t *o_car) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const uint32_t prev = s ? kpos[s - 1] : 0u;
  if (kpos[s] == prev) return;  // dropped
  const uint32_t b = base[s], e = incl[s];
  const uint32_t p = prev, nb = kinc[s] - (e - b);
  o_cls[p] = scls[s];
  o_lev[p] = slev[s];
  o_base[p] = nb;
  for (uint32_t a = b; a < e; ++a) o_car[nb + (a - b)] = rep ? dense[rep[carrier[a]]] - 1u : carrier[a];
}

// Vertical maps: the order-k contraction element (dense representative id;
// the minimum itself for order one or when the host descends) of the first
// arm's carrier of every order-k saddle, stored at the saddle's rank-index
// position, which is also its position as a minimum of order k + 1.
__global__ void k_vm_elem(int64_t S, const uint32_t *__restrict__ base, const uint32_t *__restrict__ carrier,
                          const uint32_t *__restrict__ rep, const uint32_t *__restrict__ dense, uint32_t *out) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const uint32_t c = carrier[base[s]];
  out[s] = (rep && c != kInvalid) ? dense[rep[c]] - 1u : c;
}

// Raw minimum of every dense representative id (the inverse of dense - 1).
__global__ void k_raw_of_dense(int64_t M, const uint32_t *__restrict__ dense, uint32_t *out) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const uint32_t prev = m ? dense[m - 1] : 0u;
  if (dense[m] != prev) out[prev] = (uint32_t)m;
}

// FP64 squared distance from every point to the last of its listed neighbours;
// 0 (the list is never used) when the traversal that made the list dropped a
// subtree (flag 4) or left a slot empty.
__global__ void k_nn_last_d2(int64_t n, const uint32_t *__restrict__ nn, const uint8_t *__restrict__ flags,
                             const double *__restrict__ X, const double *__restrict__ Y,
                             const double *__restrict__ Z, double *d2) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  bool whole = (flags[p] & 4) == 0;
  for (int j = 0; j < kNearK; ++j) whole &= nn[p * kNearK + j] < (uint64_t)n;
  const uint32_t q = nn[p * kNearK + kNearK - 1];
  if (!whole) {
    d2[p] = 0.0;
    return;
  }
  const double dx = X[q] - X[p], dy = Y[q] - Y[p], dz = Z[q] - Z[p];
  d2[p] = dx * dx + dy * dy + dz * dz;
}

__global__ void k_gather_u32(const uint32_t *__restrict__ src, const uint32_t *__restrict__ idx, int64_t n,
                             uint32_t *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

}  // namespace

void tower_vm_gather(Context &c, const std::vector<uint32_t> &pos, std::vector<uint32_t> &out) {
  TowerWork &R = *c.tower_work;
  const int64_t q = (int64_t)pos.size();
  out.assign(q, kInvalid);
  if (!q) return;
  if (!R.vm_stream) MHGP_HIP(hipStreamCreateWithFlags(&R.vm_stream, hipStreamNonBlocking));
  hipStream_t s = R.vm_stream;
  R.vm_idx.ensure(q);
  R.vm_out.ensure(q);
  MHGP_HIP(hipMemcpyAsync(R.vm_idx.p, pos.data(), q * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_gather_u32, dim3(blocks_for(q, 256)), dim3(256), 0, s, R.vm_elem.p, R.vm_idx.p, q, R.vm_out.p);
  MHGP_LAUNCH_CHECK();
  copy_wait(out.data(), R.vm_out.p, q * 4, hipMemcpyDeviceToHost, s);
}

void tower_vm_store(Context &c, hipStream_t s, uint32_t first_pos, const std::vector<uint32_t> &elems) {
  if (elems.empty()) return;
  copy_wait(c.tower_work->vm_elem.p + first_pos, elems.data(), elems.size() * 4, hipMemcpyHostToDevice, s);
}

void tower_prefetch_neighbours(Context &c) {
  if (c.n <= kNearK) return;
  if (!c.tower_work) c.tower_work = std::make_unique<TowerWork>();
  TowerWork &W = *c.tower_work;
  if (!W.vm_stream) MHGP_HIP(hipStreamCreateWithFlags(&W.vm_stream, hipStreamNonBlocking));
  if (!W.nn_start) {
    MHGP_HIP(hipEventCreateWithFlags(&W.nn_start, hipEventDisableTiming));
    MHGP_HIP(hipEventCreateWithFlags(&W.nn_done, hipEventDisableTiming));
  }
  W.nn.ensure((size_t)c.n * kNearK);
  W.nn_d2.ensure(c.n);
  W.nn_flags.ensure(c.n);
  // the side stream starts after the context stream's work so far (the LBVH)
  MHGP_HIP(hipEventRecord(W.nn_start, c.stream));
  MHGP_HIP(hipStreamWaitEvent(W.vm_stream, W.nn_start, 0));
  knn_all(c, kNearK, W.nn.p, W.nn_flags.p, W.vm_stream);
  hipLaunchKernelGGL(k_nn_last_d2, dim3(blocks_for(c.n, 256)), dim3(256), 0, W.vm_stream, c.n, W.nn.p, W.nn_flags.p,
                     c.X.p, c.Y.p, c.Z.p, W.nn_d2.p);
  MHGP_LAUNCH_CHECK();
  MHGP_HIP(hipEventRecord(W.nn_done, W.vm_stream));
  W.nn_pending = true;
}

void tower_rank_index(Context &c) {
  CatalogBuffers &B = *c.catalog;
  hipStream_t s = c.stream;
  const int64_t n = B.count;
  if (!c.tower_work) c.tower_work = std::make_unique<TowerWork>();
  TowerWork &W = *c.tower_work;
  std::fill(W.rank_lo, W.rank_lo + 256, 0u);
  std::fill(W.rank_hi, W.rank_hi + 256, 0u);
  if (n == 0) return;
  W.rank_keys.ensure(n);
  W.rank_vals.ensure(n);
  W.vm_elem.ensure(n);
  // neighbour lists of the points, for the descents' ball queries: started
  // beside the level order, or computed here
  W.nn_ready = false;
  if (c.n > kNearK) {
    if (!W.nn_pending) tower_prefetch_neighbours(c);
    MHGP_HIP(hipStreamWaitEvent(s, W.nn_done, 0));
    W.nn_pending = false;
    W.nn_ready = true;
  }
  W.rank_bounds.ensure(512);
  hipLaunchKernelGGL(k_rank_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, B.events.p, B.order.p, n,
                     W.rank_keys.p, W.rank_vals.p);
  c.sorter.sort(W.rank_keys.p, W.rank_vals.p, n, 0, 8, s);
  hipLaunchKernelGGL(k_fill_u32, dim3(2), dim3(256), 0, s, W.rank_bounds.p, (int64_t)512, 0u);
  hipLaunchKernelGGL(k_rank_bounds, dim3(blocks_for(n, 256)), dim3(256), 0, s, W.rank_keys.p, n, W.rank_bounds.p,
                     W.rank_bounds.p + 256);
  uint32_t h[512];
  copy_wait(h, W.rank_bounds.p, sizeof h, hipMemcpyDeviceToHost, s);
  std::copy(h, h + 256, W.rank_lo);
  std::copy(h + 256, h + 512, W.rank_hi);
  for (int r = 0; r < 256; ++r)
    if (W.rank_lo[r] > W.rank_hi[r] || W.rank_hi[r] > (uint64_t)n) throw std::runtime_error("malformed rank index");
}

void tower_order_device(Context &c, TowerWork &W, const TowerWork &R, int k, TowerOrderData &out) {
  CatalogBuffers &B = *c.catalog;
  hipStream_t s = W.stream;
  out = TowerOrderData{};
  // host clock at the synchronisations that end each phase (MHGP_TIMING)
  using clk = std::chrono::steady_clock;
  clk::time_point ph[8];
  for (auto &x : ph) x = clk::now();
  // minima of order k (k >= 2) and saddles of order k: records of closed
  // rank k and k + 1, as runs of the rank index
  const uint32_t M = R.rank_hi[k] - R.rank_lo[k], S = R.rank_hi[k + 1] - R.rank_lo[k + 1];
  const uint32_t *minima_pos = R.rank_vals.p + R.rank_lo[k], *saddle_pos = R.rank_vals.p + R.rank_lo[k + 1];
  out.minima = M;
  uint64_t cap = 1;
  while (cap < 2ull * M + 16) cap <<= 1;
  auto &tkeys = W.tkeys;
  auto &tvals = W.tvals, &labels = W.labels;
  tkeys.ensure(cap);
  tvals.ensure(cap);
  labels.ensure((size_t)M * k + 1);
  MHGP_HIP(hipMemsetAsync(tkeys.p, 0, cap * 8, s));
  if (M)
    hipLaunchKernelGGL(k_insert_minima, dim3(blocks_for(M, 256)), dim3(256), 0, s, B.events.p, B.order.p,
                       minima_pos, (int64_t)M, k, labels.p, tkeys.p, tvals.p, cap - 1);
  out.saddles = S;
  if (S == 0) return;
  auto &sizes = W.sizes, &carrier = W.carrier, &scls = W.scls;
  auto &slev = W.slev;
  auto &steps = W.steps;
  sizes.ensure(S);
  hipLaunchKernelGGL(k_arm_sizes, dim3(blocks_for(S, 256)), dim3(256), 0, s, B.events.p, B.order.p,
                     saddle_pos, (int64_t)S, sizes.p);
  W.scanner.inclusive(sizes.p, S, s);
  uint32_t arms = 0;
  copy_wait(&arms, sizes.p + S - 1, 4, hipMemcpyDeviceToHost, s);
  ph[1] = ph[2] = ph[3] = clk::now();
  // exclusive bases: shift the inclusive scan by one
  auto &base = W.base;
  base.ensure(S);
  MHGP_HIP(hipMemsetAsync(base.p, 0, 4, s));
  if (S > 1) MHGP_HIP(hipMemcpyAsync(base.p + 1, sizes.p, (S - 1) * 4, hipMemcpyDeviceToDevice, s));
  carrier.ensure(arms);
  auto &arm_saddle = W.arm_saddle;
  arm_saddle.ensure(arms);
  hipLaunchKernelGGL(k_arm_owner, dim3(blocks_for(S, 256)), dim3(256), 0, s, base.p, sizes.p, (int64_t)S,
                     arm_saddle.p);
  steps.ensure(1);
  MHGP_HIP(hipMemsetAsync(steps.p, 0, 8, s));
  ArmParams A{};
  A.ev = B.events.p; A.order = B.order.p; A.saddle_pos = saddle_pos; A.arm_base = base.p;
  A.arm_saddle = arm_saddle.p; A.arms = arms; A.k = k;
  A.table_keys = tkeys.p; A.table_vals = tvals.p; A.mask = cap - 1; A.labels = labels.p;
  A.X = c.X.p; A.Y = c.Y.p; A.Z = c.Z.p;
  A.wide = c.wide.p; A.LX = c.LX.p; A.LY = c.LY.p; A.LZ = c.LZ.p; A.leaf_id = c.leaf_id.p; A.leaf_pos = c.leaf_pos.p; A.n = c.n;
  A.carrier = carrier.p;
  A.steps = (unsigned long long *)steps.p;
  A.nn = R.nn_ready ? R.nn.p : nullptr;
  A.nn_d2 = R.nn_ready ? R.nn_d2.p : nullptr;
  auto &queue = W.queue;
  auto &qcount = W.qcount;
  queue.ensure(arms);
  qcount.ensure(1);
  MHGP_HIP(hipMemsetAsync(qcount.p, 0, 8, s));
  W.clock.start(s);
  auto &qkey = W.qkey;
  qkey.ensure(arms);
  // distinct facets of the arms to descend (load factor below 1/2)
  uint64_t fcap = 1;
  while (fcap < 2ull * arms + 16) fcap <<= 1;
  W.facet_slots.ensure(fcap);
  W.alias_arm.ensure(arms);
  W.alias_rep.ensure(arms);
  W.alias_count.ensure(1);
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for((int64_t)fcap, 256)), dim3(256), 0, s, W.facet_slots.p,
                     (int64_t)fcap, 0ull);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, s, W.alias_count.p, (int64_t)1, 0ull);
  const FacetSet D{(unsigned long long *)W.facet_slots.p, fcap - 1, W.alias_arm.p, W.alias_rep.p,
                   (unsigned long long *)W.alias_count.p};
  hipLaunchKernelGGL(k_arm_lookup, dim3(blocks_for(arms, 256)), dim3(256), 0, s, A, D, queue.p, qkey.p,
                     (unsigned long long *)qcount.p);
  uint64_t nq = 0, na = 0;
  MHGP_HIP(hipMemcpyAsync(&nq, qcount.p, 8, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipMemcpyAsync(&na, W.alias_count.p, 8, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipStreamSynchronize(s));
  ph[2] = ph[3] = clk::now();
  if (nq > 1) {  // spatial order: neighbouring lanes walk neighbouring nodes
    int bits = 8;
    while (bits < 32 && (1ll << bits) < c.n) bits += 8;
    W.sorter.sort(qkey.p, queue.p, (int64_t)nq, 0, bits, s);
  }
  uint64_t steps_total = 0;
  if (nq) {
    DescentItems items[2];
    for (int q = 0; q < 2; ++q) {
      W.dF[q].ensure((size_t)kMaxLabel * nq);
      W.dH[q].ensure(4 * nq);
      W.dO[q].ensure(nq);
      items[q] = DescentItems{W.dF[q].p, W.dH[q].p, W.dO[q].p, nq};
    }
    W.dcount.ensure(1);
    uint32_t active = (uint32_t)nq;
    for (int round = 0; active > 0; ++round) {
      const int q = round & 1;
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, W.dcount.p, (int64_t)1, 0ull);
      hipLaunchKernelGGL(k_descend_round, dim3(blocks_for(active, kArmBlock)), dim3(kArmBlock), 0, s, A, round,
                         queue.p, items[q ^ 1], active, items[q], (unsigned long long *)W.dcount.p);
      uint64_t nxt = 0;
      copy_wait(&nxt, W.dcount.p, 8, hipMemcpyDeviceToHost, s);
      steps_total += nxt;
      active = (uint32_t)nxt;
    }
    ph[3] = clk::now();
  }
  if (na)
    hipLaunchKernelGGL(k_alias_carriers, dim3(blocks_for(na, 256)), dim3(256), 0, s, W.alias_arm.p, W.alias_rep.p,
                       na, carrier.p);
  W.clock.stop(s);
  MHGP_LAUNCH_CHECK();
  W.arm_ms += W.clock.ms();
  W.arm_launches++;
  // algorithmic bytes: each saddle record and its order/position entries read
  // once, each arm's carrier written once and one hash probe (key + value)
  W.arm_bytes += (int64_t)S * ((int64_t)sizeof(EventRec) + 8) + (int64_t)arms * (4 + 12);
  scls.ensure(S);
  slev.ensure(S);
  hipLaunchKernelGGL(k_saddle_meta, dim3(blocks_for(S, 256)), dim3(256), 0, s, B.events.p, B.order.p, B.cls.p,
                     saddle_pos, (int64_t)S, scls.p, slev.p);
  // drop the saddles that cannot change the contraction (unless the host
  // must descend some arms: then every saddle goes down as it is)
  const int64_t minima_n = k == 1 ? c.n : (int64_t)M;
  auto &first = W.first, &keep = W.keep, &kept = W.kept, &flag1 = W.flag1;
  first.ensure(minima_n);
  keep.ensure(S);
  kept.ensure(S);
  flag1.ensure(1);
  MHGP_HIP(hipMemsetAsync(flag1.p, 0, 4, s));
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(minima_n, 256)), dim3(256), 0, s, first.p, minima_n, kInvalid);
  hipLaunchKernelGGL(k_first_touch, dim3(blocks_for(arms, 256)), dim3(256), 0, s, arm_saddle.p, carrier.p,
                     (int64_t)arms, minima_n, first.p, flag1.p);
  uint32_t any_invalid = 0;
  copy_wait(&any_invalid, flag1.p, 4, hipMemcpyDeviceToHost, s);
  ph[4] = ph[5] = ph[6] = clk::now();
  uint32_t S2 = S, arms2 = arms;
  const uint32_t *d_cls = scls.p, *d_base = base.p, *d_car = carrier.p;
  const double *d_lev = slev.p;
  auto &c_cls = W.c_cls, &c_base = W.c_base, &c_car = W.c_car;
  auto &c_lev = W.c_lev;
  uint32_t msf_edges = 0;
  int msf_rounds = 0;
  if (!any_invalid) {
    const dim3 tb(256);
    const int all_owned = k == 1 ? 1 : 0;
    const dim3 gm(blocks_for(minima_n, 256)), gs(blocks_for(S, 256)), ga(blocks_for(arms, 256));
    // attachment forest of the fresh minima and its representatives
    auto &attach = W.attach, &rep = W.rep, &dense = W.dense, &err = W.err;
    attach.ensure(minima_n);
    rep.ensure(minima_n);
    err.ensure(1);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), tb, 0, s, err.p, (int64_t)1, 0u);
    hipLaunchKernelGGL(k_fill_u32, gm, tb, 0, s, attach.p, minima_n, kInvalid);
    hipLaunchKernelGGL(k_attach, gs, tb, 0, s, (int64_t)S, base.p, sizes.p, carrier.p, first.p, all_owned, attach.p);
    hipLaunchKernelGGL(k_attach_roots, gm, tb, 0, s, minima_n, attach.p, rep.p, err.p);
    // star edges between the representatives of owned carriers, in arm (= level) order
    auto &eflag = W.eflag, &eu = W.eu, &ev = W.ev, &es = W.es;
    eflag.ensure(arms);
    hipLaunchKernelGGL(k_edge_flags, ga, tb, 0, s, arm_saddle.p, base.p, sizes.p, carrier.p, first.p, rep.p,
                       all_owned, (int64_t)arms, eflag.p);
    W.scanner.inclusive(eflag.p, arms, s);
    uint32_t E = 0, bad = 0;
    MHGP_HIP(hipMemcpyAsync(&bad, err.p, 4, hipMemcpyDeviceToHost, s));
    copy_wait(&E, eflag.p + arms - 1, 4, hipMemcpyDeviceToHost, s);
    if (bad) throw std::runtime_error("tower attachment forest is cyclic");
    ph[5] = ph[6] = clk::now();
    eu.ensure(E);
    ev.ensure(E);
    es.ensure(E);
    hipLaunchKernelGGL(k_edge_list, ga, tb, 0, s, arm_saddle.p, base.p, sizes.p, carrier.p, first.p, rep.p,
                       all_owned, eflag.p, (int64_t)arms, eu.p, ev.p, es.p);
    // Boruvka spanning forest of the minima over the star edges
    auto &comp = W.comp, &comp2 = W.comp2, &best = W.best, &hook = W.hook, &msf = W.msf;
    auto &mcount = W.mcount, &changed = W.changed, &live_n = W.live_n;
    auto &Lu = W.Lu, &Lv = W.Lv, &Li = W.Li;
    for (int q = 0; q < 2; ++q) {
      Lu[q].ensure(E);
      Lv[q].ensure(E);
      Li[q].ensure(E);
    }
    live_n.ensure(1);
    comp.ensure(minima_n);
    comp2.ensure(minima_n);
    best.ensure(minima_n);
    hook.ensure(minima_n);
    msf.ensure(E + 1);
    mcount.ensure(S);
    changed.ensure(1);
    const dim3 ge(blocks_for(E + 1, 256));
    hipLaunchKernelGGL(k_identity, gm, tb, 0, s, comp.p, minima_n);
    hipLaunchKernelGGL(k_fill_u32, gm, tb, 0, s, best.p, minima_n, kInvalid);
    hipLaunchKernelGGL(k_fill_u32, gm, tb, 0, s, hook.p, minima_n, kInvalid);
    hipLaunchKernelGGL(k_fill_u32, ge, tb, 0, s, msf.p, (int64_t)E, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(S, 256)), tb, 0, s, mcount.p, (int64_t)S, 0u);
    const uint32_t *lu = eu.p, *lv = ev.p, *li = nullptr;
    uint32_t live = E;
    msf_edges = E;
    for (int round = 0; live > 0; ++round) {
      if (round > 64) throw std::runtime_error("spanning forest of the tower did not converge");
      msf_rounds = round + 1;
      const int q = round & 1;
      hipLaunchKernelGGL(k_fill_u32, dim3(1), tb, 0, s, changed.p, (int64_t)1, 0u);
      hipLaunchKernelGGL(k_fill_u32, dim3(1), tb, 0, s, live_n.p, (int64_t)1, 0u);
      hipLaunchKernelGGL(k_msf_best, dim3(blocks_for(live, 256)), tb, 0, s, (int64_t)live, lu, lv, li, comp.p,
                         best.p, Lu[q].p, Lv[q].p, Li[q].p, live_n.p);
      hipLaunchKernelGGL(k_msf_hook, gm, tb, 0, s, minima_n, eu.p, ev.p, comp.p, best.p, hook.p, msf.p, changed.p);
      hipLaunchKernelGGL(k_msf_apply, gm, tb, 0, s, minima_n, comp.p, best.p, hook.p);
      hipLaunchKernelGGL(k_msf_flatten, gm, tb, 0, s, minima_n, comp.p, comp2.p, err.p);
      std::swap(comp.p, comp2.p);
      std::swap(comp.cap, comp2.cap);
      uint32_t flags[3] = {0, 0, 0};
      MHGP_HIP(hipMemcpyAsync(&flags[0], changed.p, 4, hipMemcpyDeviceToHost, s));
      MHGP_HIP(hipMemcpyAsync(&flags[2], live_n.p, 4, hipMemcpyDeviceToHost, s));
      copy_wait(&flags[1], err.p, 4, hipMemcpyDeviceToHost, s);
      if (flags[1]) throw std::runtime_error("spanning forest of the tower: cyclic hook forest");
      if (!flags[0]) break;
      live = flags[2];
      lu = Lu[q].p;
      lv = Lv[q].p;
      li = Li[q].p;
    }
    ph[6] = clk::now();
    if (E) hipLaunchKernelGGL(k_msf_count, ge, tb, 0, s, (int64_t)E, msf.p, es.p, mcount.p);
    // keep births, multifusions and shared levels
    hipLaunchKernelGGL(k_reduce_saddles, gs, tb, 0, s, (int64_t)S, base.p, sizes.p, carrier.p, scls.p, first.p,
                       mcount.p, all_owned, keep.p, kept.p);
    const uint32_t *d_rep = nullptr, *d_dense = nullptr;
    uint32_t Mdense = (uint32_t)minima_n;
    if (k >= 2) {  // order one keeps point ids: every point is born
      dense.ensure(minima_n);
      hipLaunchKernelGGL(k_fill_u32, gm, tb, 0, s, dense.p, minima_n, 0u);
      hipLaunchKernelGGL(k_flag_used, gs, tb, 0, s, (int64_t)S, keep.p, base.p, sizes.p, carrier.p, rep.p, dense.p);
      W.scanner.inclusive(dense.p, minima_n, s);
      copy_wait(&Mdense, dense.p + minima_n - 1, 4, hipMemcpyDeviceToHost, s);
      d_rep = rep.p;
      d_dense = dense.p;
      // dense id -> raw minimum, for the vertical maps out of this order
      W.raw_dense.ensure(Mdense + 1);
      hipLaunchKernelGGL(k_raw_of_dense, gm, tb, 0, s, minima_n, dense.p, W.raw_dense.p);
      out.raw_of_dense.resize(Mdense);
      if (Mdense) copy_wait(out.raw_of_dense.data(), W.raw_dense.p, (size_t)Mdense * 4, hipMemcpyDeviceToHost, s);
    }
    hipLaunchKernelGGL(k_vm_elem, gs, tb, 0, s, (int64_t)S, base.p, carrier.p, d_rep, d_dense,
                       R.vm_elem.p + R.rank_lo[k + 1]);
    W.scanner.inclusive(keep.p, S, s);
    W.scanner.inclusive(kept.p, S, s);
    MHGP_HIP(hipMemcpyAsync(&S2, keep.p + S - 1, 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(&arms2, kept.p + S - 1, 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipStreamSynchronize(s));
    c_cls.ensure(S2);
    c_lev.ensure(S2);
    c_base.ensure(S2);
    c_car.ensure(arms2);
    hipLaunchKernelGGL(k_compact_saddles, dim3(blocks_for(S, 256)), tb, 0, s, (int64_t)S, keep.p, kept.p, base.p,
                       sizes.p, carrier.p, scls.p, slev.p, d_rep, d_dense, c_cls.p, c_lev.p, c_base.p, c_car.p);
    d_cls = c_cls.p;
    d_lev = c_lev.p;
    d_base = c_base.p;
    d_car = c_car.p;
    if (k >= 2) out.minima = Mdense;
  }
  // page-locked results: cls | arm_base | carrier | level
  const size_t o_base = (size_t)S2 * 4, o_car = o_base + ((size_t)S2 + 1) * 4;
  const size_t o_lev = (o_car + (size_t)arms2 * 4 + 7) & ~(size_t)7;
  char *h = (char *)c.tower_pinned[k].ensure(o_lev + (size_t)S2 * 8 + 8);
  uint32_t *h_cls = (uint32_t *)h, *h_base = (uint32_t *)(h + o_base), *h_car = (uint32_t *)(h + o_car);
  double *h_lev = (double *)(h + o_lev);
  if (S2) {
    MHGP_HIP(hipMemcpyAsync(h_cls, d_cls, (size_t)S2 * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(h_base, d_base, (size_t)S2 * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(h_car, d_car, (size_t)arms2 * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(h_lev, d_lev, (size_t)S2 * 8, hipMemcpyDeviceToHost, s));
  }
  const uint64_t st = steps_total;
  MHGP_HIP(hipStreamSynchronize(s));
  ph[7] = clk::now();
  h_base[S2] = arms2;
  out.saddles = S2;
  out.cls = h_cls;
  out.arm_base = h_base;
  out.carrier = h_car;
  out.level = h_lev;
  out.arms = arms2;
  out.steps = (int64_t)st;
  if (std::getenv("MHGP_TIMING"))
    std::fprintf(stderr,
                 "[mhgp] arms order %2d: %u saddles, %u arms, %llu descending, %llu steps, %u forest edges in %d "
                 "rounds, %u kept | ms: setup %.2f lookup %.2f descents %.2f first %.2f edges %.2f forest %.2f "
                 "tail %.2f\n",
                 k, S, arms, (unsigned long long)nq, (unsigned long long)st, msf_edges, msf_rounds, S2,
                 std::chrono::duration<double, std::milli>(ph[1] - ph[0]).count(),
                 std::chrono::duration<double, std::milli>(ph[2] - ph[1]).count(),
                 std::chrono::duration<double, std::milli>(ph[3] - ph[2]).count(),
                 std::chrono::duration<double, std::milli>(ph[4] - ph[3]).count(),
                 std::chrono::duration<double, std::milli>(ph[5] - ph[4]).count(),
                 std::chrono::duration<double, std::milli>(ph[6] - ph[5]).count(),
                 std::chrono::duration<double, std::milli>(ph[7] - ph[6]).count());
  const bool need_labels = any_invalid != 0;  // then the saddles were not filtered
  out.vm_first_pos = R.rank_lo[k + 1];
  if (need_labels) {
    // the host descends unresolved arms; it needs the saddle labels, their
    // supports and the minima labels
    out.minima_labels.resize((size_t)M * k);
    if (M) copy_wait(out.minima_labels.data(), labels.p, (size_t)M * k * 4, hipMemcpyDeviceToHost, s);
    auto &slab = W.slab, &ssup = W.ssup;
    slab.ensure((size_t)S * kMaxLabel);
    ssup.ensure(4 * (size_t)S);
    hipLaunchKernelGGL(k_saddle_labels, dim3(blocks_for(S, 256)), dim3(256), 0, s, B.events.p, B.order.p,
                       saddle_pos, (int64_t)S, slab.p, ssup.p);
    out.saddle_labels.resize((size_t)S * kMaxLabel);
    out.sup.resize(4 * (size_t)S);
    MHGP_HIP(hipMemcpyAsync(out.saddle_labels.data(), slab.p, (size_t)S * kMaxLabel * 4, hipMemcpyDeviceToHost, s));
    copy_wait(out.sup.data(), ssup.p, 4 * (size_t)S * 4, hipMemcpyDeviceToHost, s);
  }
}

}  // namespace mhgp
