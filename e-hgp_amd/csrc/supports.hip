// Critical-sphere catalogue (supports of 2..4 points, closed rank <= s_max).
//
// Observable definition (reference reference/morsehgp3d_oracle/catalog.py and
// morsehgp3d/src/cpu/hierarchy/critical_catalog.cpp): a support U of 1..4
// affinely independent points whose circumcentre c (in aff U) lies strictly
// inside conv U; interior I = points strictly inside the sphere, shell = points
// on it; an event when shell == U and |I| + |U| <= s_max, an extra-shell
// diagnostic when shell != U and |I| + |U| <= s_max.
//
// Enumeration (this backend's own method, replacing the reference's node-tuple
// branch-and-bound of pair_support_stream.cpp / higher_support_stream.cpp):
// every support is enumerated once, from its least PointId u (the anchor).
//  1. Direction-cell certificate.  For a critical sphere through u with centre
//     u + r d, every open ball of radius rho tangent at u towards d lies in
//     the open sphere when rho <= r.  The unit sphere of directions is cut
//     into 384 cube-map cells; for each cell C the radius rho_C is the
//     (s_max-1)-th smallest radius at which a gathered point lies strictly
//     inside every tangent ball of the cell.  A sphere with direction in C and
//     r >= rho_C therefore holds >= s_max-1 interior points and is not
//     critical.  Since the centre of a well-centred support lies in the
//     cloud's AABB, r is also below the AABB exit distance of the cell.
//  2. Region.  All points of a critical sphere with direction in C lie in
//     the union of closed tangent balls of radius rho'_C over C; a point's
//     mask records the cells whose union contains it.  Pairs, triangles and
//     tetrahedra through u are searched among region members only, and their
//     interior is counted over the whole region.
//  3. Tetrahedra come from faces (u, v, w) (the three least ids) by the
//     pencil of spheres through the face's circumcircle: the apex z must be
//     among the first s_max-3 points swept on its side, so only those are
//     tested.
// Every emitted record is certified with the FP64 predicates of
// predicates.hpp over the anchor's region; anything undecided goes to the
// host fallback queue with the anchor's region.
#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>

#include "batch.hpp"
#include "wave.hpp"
#include "pipeline.hpp"
#include "predicates.hpp"
#include "supports.hpp"
#include "timing.hpp"
#include "traverse.hpp"

namespace mhgp {

int64_t catalog_device_bytes(const CatalogBuffers &b) {
  return b.radius.bytes() + b.gcount.bytes() + b.goff.bytes() + b.glist.bytes() + b.need.bytes() +
         b.status.bytes() + b.events.bytes() + b.fallbacks.bytes() + b.fb_region.bytes() +
         b.counters.bytes() + b.cells.bytes() + b.cellf.bytes() + b.scratch.bytes() + b.anchors.bytes() + b.overflow.bytes() +
         b.total.bytes() + b.run_count.bytes() + b.run_pos.bytes() + b.run_gidx.bytes() + b.run_new.bytes() +
         b.run_link.bytes() + b.run_rec.bytes();
}

namespace {

constexpr int kBlock = 256;

// LDS written by some lanes of a wave, then read by other lanes of it
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

enum : uint32_t { ST_PENDING = 0, ST_DONE = 1, ST_WIDEN = 2, ST_BIG = 3, ST_HUGE = 4, ST_PROBED = 5 };

// Regions of the global-memory path are enumerated face by face in
// O(region^3) work; beyond the context's region limit (at most this) one
// anchor would keep a workgroup busy for many seconds.  Such an anchor is not
// enumerated: the catalogue is returned complete for every other anchor with
// status conditional (a budget stop), and the anchor and its region size are
// reported.
constexpr uint32_t kBigRegionMax = kRegionLimitMax;

struct Params {
  const double *X, *Y, *Z;       // canonical coordinates
  const uint32_t *gcount;
  const uint64_t *goff;
  const uint32_t *glist;
  const double *radius;
  const double *cells;           // 8 doubles per cell
  const float *cellf;            // 8 floats per cell: the FP32 cell (see cell_f)
  double blo[3], bhi[3];         // cloud AABB
  int s_max;
  double *need;
  uint32_t *status;
  EventRec *events;
  uint64_t event_cap;
  FallbackRec *fallbacks;
  uint64_t fallback_cap;
  uint32_t *fb_region;
  uint64_t fb_region_cap;
  unsigned long long *counters;  // [0] events [1] fallbacks [2] region words [3] candidates
  const uint32_t *anchors;
  uint32_t anchor_count;
  int big;                       // 1: arrays live in global scratch slabs
  int probe;                     // 1: stop once the region size is known (ST_PROBED)
  double *scratch;
  uint64_t slab_doubles;
  uint32_t gcap, rcap;
  int count_flops;               // 1: count the FP64 (counters[5]) and FP32 (counters[6]) operations
  uint32_t region_limit;         // largest region the global-memory path enumerates
};

struct Work {
  double *gx, *gy, *gz;
  uint32_t *gid;
  float *rho;        // certificate radius per cell (FP32, rounded up)
  uint32_t *ridx;    // gathered index of each region slot (until the points are compacted)
  uint64_t *rmask;  // 6 words per member
  uint32_t *kidx;
  float4 *fc;       // FP32 offset and distance per gathered point (certificate and masks)
  uint32_t *cand;   // candidate supports classified (LDS, flushed once per anchor)
  unsigned long long *flops;  // FP64 operations of the certified predicates (LDS, flushed once per anchor; null: not counted)
  float *fx, *fy, *fz;        // FP32 copies of the region's coordinates (LDS path only)
  unsigned long long *flops32;  // FP32 operations of the tests (as flops; null: not counted)
};

// FP64 operations (add, subtract, multiply; a fused multiply-add counts two)
// of one evaluation of each certified predicate, counted from predicates.hpp:
// the value, its magnitude bound and the certification test.
constexpr uint32_t kFlopPairSide = 18;   // PairSphere::side
constexpr uint32_t kFlopTetSide = 21;    // TetSphere::side
constexpr uint32_t kFlopCapsule = 8;     // Capsule::outside
constexpr uint32_t kFlopFacePoint = 33;  // TriSphere::power + height, two certified signs
constexpr uint32_t kFlopSweepKey = 24;   // sweep_key
constexpr uint32_t kFlopTriInit = 66;    // TriSphere::init
constexpr uint32_t kFlopTetInit = 150;   // TetSphere::init

// Counting is off unless the launch asks for it (flops == nullptr): the
// timed path pays for one uniform branch per call.
__device__ inline void add_flops(const Work &w, uint64_t v) {
  if (w.flops && v) atomicAdd(w.flops, (unsigned long long)v);
}
// FP32 operations of the FP32 tests, counted the same way (the certificate
// and mask tests of a cell, the capsule test, the face, pair and tetrahedron
// polynomials with their bounds, a key interval); reciprocals and square
// roots count one each.
constexpr uint32_t kF32Cover = 16;      // one point against one cell (certificate)
constexpr uint32_t kF32Mask = 16;       // one point against one cell (membership)
constexpr uint32_t kF32Capsule = 8;     // capsule_chunk_f, per slot
constexpr uint32_t kF32FacePoint = 31;  // TriSphereF::eval and two certified signs
constexpr uint32_t kF32Key = 16;        // key_bounds_32
constexpr uint32_t kF32PairSide = 17;   // PairSphereF::side
constexpr uint32_t kF32TetSide = 20;    // TetSphereF::side
__device__ inline void add_flops32(const Work &w, uint64_t v) {
  if (w.flops32 && v) atomicAdd(w.flops32, (unsigned long long)v);
}

__device__ inline V3 gpt(const Work &w, uint32_t i) { return {w.gx[i], w.gy[i], w.gz[i]}; }

// Interior id lists are written and sorted with static indices only, so
// they stay in registers: an append is a select over the slots, the sort an
// odd-even transposition network (unused slots hold kInvalid and sort last).
__device__ inline void put_static(uint32_t (&a)[kMaxInterior], int pos, uint32_t v) {
#pragma unroll
  for (int j = 0; j < kMaxInterior; ++j) a[j] = j == pos ? v : a[j];
}

__device__ inline void sort_static(uint32_t (&a)[kMaxInterior], int n) {
#pragma unroll
  for (int j = 0; j < kMaxInterior; ++j) a[j] = j < n ? a[j] : kInvalid;
#pragma unroll
  for (int pass = 0; pass < kMaxInterior; ++pass) {
#pragma unroll
    for (int j = pass & 1; j + 1 < kMaxInterior; j += 2) {
      const uint32_t x = a[j], y = a[j + 1];
      a[j] = min(x, y);
      a[j + 1] = max(x, y);
    }
  }
}

__device__ void emit_event(const Params &P, uint32_t u, const uint32_t *sup, int m, int nint,
                           const uint32_t *interior, double level, double lerr) {
  const unsigned long long slot = wave_append(true, &P.counters[0]);
  if (slot >= P.event_cap) return;
  EventRec &e = P.events[slot];
  e.ids[0] = u;
  for (int i = 1; i < 4; ++i) e.ids[i] = i < m ? sup[i - 1] : kInvalid;
  e.meta = (uint32_t)m | ((uint32_t)(nint + m) << 8) | ((uint32_t)nint << 16);
#pragma unroll
  for (int i = 0; i < kMaxInterior; ++i) e.interior[i] = i < nint ? interior[i] : kInvalid;
  e.level = level;
  // rounded up into float (an infinite bound only sends the record to the
  // exact host ordering)
  e.level_err = (float)lerr * (1.0f + 1e-6f);
}

// Queues undecided work and, once per anchor, its region for the host.
__device__ void emit_fallback(const Params &P, const Work &w, uint32_t u, uint32_t kind,
                              const uint32_t *ids, int m, uint32_t M, uint32_t *region_slot) {
  if (*region_slot == kInvalid) {
    uint32_t expected = kInvalid;
    // first fallback of this anchor publishes its region
    unsigned long long off = atomicAdd(&P.counters[2], (unsigned long long)M);
    if (off + M <= P.fb_region_cap) {
      for (uint32_t s = 0; s < M; ++s) P.fb_region[off + s] = w.gid[s];
      atomicCAS(region_slot, expected, (uint32_t)off);
    }
  }
  unsigned long long slot = atomicAdd(&P.counters[1], 1ull);
  if (slot >= P.fallback_cap) return;
  FallbackRec &f = P.fallbacks[slot];
  f.kind = kind;
  f.ids[0] = u;
  for (int i = 1; i < 4; ++i) f.ids[i] = i < m ? ids[i - 1] : kInvalid;
  f.region_off = *region_slot;
  f.region_len = M;
  f.pad = 0;
}

// Certified classification of the region against a candidate sphere.
// Returns 1 event, 0 not critical, -1 undecided (host).
// FP32 certification of a sign: v against its magnitude bound a.  With
// the coordinates, the coefficients and the at most ~9 roundings of a chain
// in FP32 (the FP64 coefficients rounded to nearest), |v32 - v| < 1e-6 a
// for the polynomials below; a sign is certified when |v32| > 1e-5 a and
// a > 1e-30 (FP32 underflow then adds at most ~1e-37, far below the
// threshold).  0 leaves the point to the FP64 test.
__device__ inline int cert32(float v, float a) {
  const float e = 1e-5f * a;
  return !(a > 1e-30f) ? 0 : (v > e ? 1 : (v < -e ? -1 : 0));
}

// Pair and tetrahedron side tests in FP32 on the region's FP32 coordinates.
struct PairSphereF {
  float a[3];
  __device__ int side(float x, float y, float z) const {
    const float v = x * (x - a[0]) + y * (y - a[1]) + z * (z - a[2]);
    const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
    const float va = ax * (ax + fabsf(a[0])) + ay * (ay + fabsf(a[1])) + az * (az + fabsf(a[2]));
    return cert32(v, va);
  }
};
struct TetSphereF {
  float V, VA, N[3], NA[3];
  int sg;  // sign of the FP64 volume (certified nonzero)
  __device__ int side(float x, float y, float z) const {
    const float pp = x * x + y * y + z * z;
    const float Q = V * pp - (x * N[0] + y * N[1] + z * N[2]);
    const float QA = VA * pp + (fabsf(x) * NA[0] + fabsf(y) * NA[1] + fabsf(z) * NA[2]);
    return sg * cert32(Q, QA);
  }
};
__device__ inline PairSphereF to_f32(const PairSphere &s) {
  return {{(float)s.a.x, (float)s.a.y, (float)s.a.z}};
}
__device__ inline TetSphereF to_f32(const TetSphere &t) {
  return {(float)t.V, (float)t.VA, {(float)t.N.x, (float)t.N.y, (float)t.N.z},
          {(float)t.NA.x, (float)t.NA.y, (float)t.NA.z}, t.V > 0 ? 1 : -1};
}

template <class S>
__device__ int classify(const Params &P, const Work &w, uint32_t M, const S &sph,
                        const uint32_t *skip, int m, uint32_t (&interior)[kMaxInterior], int &nint) {
  nint = 0;
  const int cap = P.s_max - m;
  int undecided = 0;
  constexpr uint32_t per = sizeof(S) == sizeof(PairSphere) ? kFlopPairSide : kFlopTetSide;
  constexpr uint32_t per32 = sizeof(S) == sizeof(PairSphere) ? kF32PairSide : kF32TetSide;
  uint32_t evals = 0, evals32 = 0;
  // FP32 first where the region has FP32 copies; FP64 for what it leaves open
  const bool f32 = w.fx != nullptr;
  const auto s32 = to_f32(sph);
  for (uint32_t s = 0; s < M; ++s) {
    const uint32_t gi = s;
    bool on_support = false;
    for (int k = 0; k < m - 1; ++k) on_support |= (gi == skip[k]);
    if (on_support) continue;
    int side = f32 ? s32.side(w.fx[gi], w.fy[gi], w.fz[gi]) : 0;
    evals32 += f32;
    if (side == 0) {
      side = sph.side(gpt(w, gi));
      evals++;
    }
    if (side < 0) {
      put_static(interior, nint, w.gid[gi]);
      nint++;
      if (nint > cap) {
        add_flops(w, (uint64_t)evals * per);
        add_flops32(w, (uint64_t)evals32 * per32);
        return 0;
      }
    } else if (side == 0) {
      undecided++;
    }
  }
  add_flops(w, (uint64_t)evals * per);
  add_flops32(w, (uint64_t)evals32 * per32);
  return undecided ? -1 : 1;
}


__device__ void process_pair(const Params &P, const Work &w, uint32_t u, uint32_t M, uint32_t vi,
                             uint32_t *region_slot) {
  PairSphere sph;
  sph.init(gpt(w, vi));
  uint32_t interior[kMaxInterior];
  int nint;
  uint32_t skip[1] = {vi};
  int r = classify(P, w, M, sph, skip, 2, interior, nint);
  uint32_t vid = w.gid[vi];
  if (r == 0) return;
  atomicAdd(w.cand, 1u);
  if (r < 0) {
    emit_fallback(P, w, u, 2, &vid, 2, M, region_slot);
    return;
  }
  sort_static(interior, nint);
  emit_event(P, u, &vid, 2, nint, interior, sph.level(), 8 * kCert);
}

__device__ void process_tet(const Params &P, const Work &w, uint32_t u, uint32_t M, uint32_t vi, uint32_t wi,
                            uint32_t zi, uint32_t *region_slot) {
  TetSphere t;
  int wc = t.init(gpt(w, vi), gpt(w, wi), gpt(w, zi));
  add_flops(w, kFlopTetInit);
  uint32_t ids[3] = {w.gid[vi], w.gid[wi], w.gid[zi]};
  if (wc < 0) return;
  atomicAdd(w.cand, 1u);
  if (wc == 0) { emit_fallback(P, w, u, 4, ids, 4, M, region_slot); return; }
  uint32_t interior[kMaxInterior];
  int nint;
  uint32_t skip[3] = {vi, wi, zi};
  int r = classify(P, w, M, t, skip, 4, interior, nint);
  if (r == 0) return;
  if (r < 0) { emit_fallback(P, w, u, 4, ids, 4, M, region_slot); return; }
  sort_static(interior, nint);
  emit_event(P, u, ids, 4, nint, interior, t.level(), t.level_err());
}

// Smallest sweep keys of the outside points ahead of one side: tau with its
// error bound and the region slot.  Insertion runs over static indices, so
// the list stays in registers.
struct SweepList {
  static constexpr int kCap = 8;  // >= s_max - 3 for s_max <= 11
  double hi[kCap];                // upper key bounds, ascending
  float err[kCap];                // key error bounds (key = hi - err)
  uint32_t slot[kCap];
  double evicted_lo;  // smallest lower key bound of a point not in the list
  __device__ void init() {
#pragma unroll
    for (int j = 0; j < kCap; ++j) { hi[j] = INFINITY; err[j] = 0.f; slot[j] = kInvalid; }
    evicted_lo = INFINITY;
  }
  __device__ void insert(double t, double e, uint32_t s) {
    const double h = t + e;
    if (!(h < hi[kCap - 1])) {
      evicted_lo = fmin(evicted_lo, t - e);
      return;
    }
    evicted_lo = fmin(evicted_lo, hi[kCap - 1] - 2.0 * (double)err[kCap - 1]);
    const float ef = (float)e * (1.0f + 1e-6f);
#pragma unroll
    for (int j = kCap - 1; j > 0; --j) {
      const bool shift = h < hi[j - 1];
      const bool place = !shift && h < hi[j];
      hi[j] = shift ? hi[j - 1] : (place ? h : hi[j]);
      err[j] = shift ? err[j - 1] : (place ? ef : err[j]);
      slot[j] = shift ? slot[j - 1] : (place ? s : slot[j]);
    }
    if (h < hi[0]) { hi[0] = h; err[0] = ef; slot[0] = s; }
  }
  // upper key bound of the need-th entry (need in 1..kCap)
  __device__ double threshold(int need) const {
    double r = INFINITY;
#pragma unroll
    for (int j = 0; j < kCap; ++j) r = (j == need - 1) ? hi[j] : r;
    return r;
  }
};

// Largest lower key bounds of the inside points behind one side (a subset
// of them only weakens the apex pruning bound, never invalidates it).
struct BehindList {
  static constexpr int kCap = 4;
  double v[kCap];
  __device__ void init() {
#pragma unroll
    for (int j = 0; j < kCap; ++j) v[j] = -INFINITY;
  }
  __device__ void insert(double x) {
#pragma unroll
    for (int j = kCap - 1; j > 0; --j) v[j] = x > v[j - 1] ? v[j - 1] : (x > v[j] ? x : v[j]);
    v[0] = fmax(v[0], x);
  }
  __device__ int count_above(double x) const {
    int c = 0;
#pragma unroll
    for (int j = 0; j < kCap; ++j) c += v[j] > x;
    return c;
  }
};

// The same two lists in FP32, filled for both sides during a face's first
// pass: key bounds rounded outwards (lo down, hi up), so that every
// threshold test below stays conservative (wider intervals only admit more
// apex candidates, which are classified exactly).
//
// An entry keeps its upper bound and one word with the region slot (low 11
// bits, regions of up to 2048 points) and the interval's width, rounded up
// to a 21-bit float (the FP32 width's bit pattern rounded up past its 10
// lowest bits); its lower bound is
// hi - width, rounded down.  Listed entries have finite upper bounds (an
// infinite one never enters the list).  The upper bounds stay sorted with
// one median-of-three per entry.
__device__ inline uint32_t pack_entry(float h, float l, uint32_t slot) {
  // the rounded difference widened past its rounding error
  float gap = (h - l) * (1.0f + 2.4e-7f) + 1e-37f;
  if (!(gap >= 0.0f)) gap = INFINITY;
  return (((__float_as_uint(gap) + 1023u) >> 10) << 11) | slot;
}
__device__ inline float entry_lo(float h, uint32_t e) {
  const float d = h - __uint_as_float((e >> 11) << 10);
  return d - fabsf(d) * 2.4e-7f - 1e-37f;
}
__device__ inline uint32_t entry_slot(uint32_t e) { return e & 2047u; }

struct SweepListF {
  static constexpr int kCap = 8;  // >= s_max - 3 for s_max <= 11
  float hi[kCap];                 // upper key bounds, ascending
  uint32_t ent[kCap];             // slot and width (pack_entry); kInvalid: empty
  float evicted_lo;  // smallest lower key bound of a point not in the list
  __device__ void init() {
#pragma unroll
    for (int j = 0; j < kCap; ++j) { hi[j] = INFINITY; ent[j] = kInvalid; }
    evicted_lo = INFINITY;
  }
  __device__ void insert(float h, float l, uint32_t s) {
    if (!(h < hi[kCap - 1])) {
      evicted_lo = fminf(evicted_lo, l);
      return;
    }
    if (ent[kCap - 1] != kInvalid) evicted_lo = fminf(evicted_lo, entry_lo(hi[kCap - 1], ent[kCap - 1]));
    const uint32_t e = pack_entry(h, l, s);
#pragma unroll
    for (int j = kCap - 1; j > 0; --j) {
      ent[j] = h < hi[j - 1] ? ent[j - 1] : (h < hi[j] ? e : ent[j]);
      hi[j] = __builtin_amdgcn_fmed3f(hi[j - 1], h, hi[j]);
    }
    ent[0] = h < hi[0] ? e : ent[0];
    hi[0] = fminf(hi[0], h);
  }
  __device__ float threshold(int need) const {
    float r = INFINITY;
#pragma unroll
    for (int j = 0; j < kCap; ++j) r = (j == need - 1) ? hi[j] : r;
    return r;
  }
};
struct BehindListF {
  static constexpr int kCap = 4;
  float v[kCap];
  __device__ void init() {
#pragma unroll
    for (int j = 0; j < kCap; ++j) v[j] = -INFINITY;
  }
  __device__ void insert(float x) {
    // descending: the median of the entry above, x and the entry
#pragma unroll
    for (int j = kCap - 1; j > 0; --j) v[j] = __builtin_amdgcn_fmed3f(v[j - 1], x, v[j]);
    v[0] = fmaxf(v[0], x);
  }
  __device__ int count_above(float x) const {
    int c = 0;
#pragma unroll
    for (int j = 0; j < kCap; ++j) c += v[j] > x;
    return c;
  }
};

// Sweep key of a point against the face (tau = P / |H|) and its error bound.
// The reciprocal of |H| is the hardware estimate refined by two Newton
// steps; the residual e = 1 - |H| r of the result bounds its relative error
// (up to the rounding of e), which joins the key's error bound.  Outside the
// range where the estimate is reliable the key is unknown: tau = 0 with an
// infinite bound, which meets every threshold and sends the side to its full
// scan.
__device__ inline void sweep_key(double Pw, double PA, double H, double HA, double &tau, double &terr) {
  const double h = fabs(H);
  double r = __builtin_amdgcn_rcp(h);
  double e = __builtin_fma(-h, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-h, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double rerr = 2.0 * fabs(__builtin_fma(-h, r, 1.0));
  tau = Pw * r;
  terr = (kCert * PA + kCertAbs + (kCert * HA + kCertAbs) * fabs(tau)) * r * (1.0 + 1e-12) + (1e-15 + rerr) * fabs(tau);
  if (!(h > 1e-290 && h < 1e290)) {
    tau = 0.0;
    terr = INFINITY;
  }
}

// FP32 bounds of the key interval [tau - terr, tau + terr], rounded outwards.
__device__ inline void key_bounds_f(double tau, double terr, float &hf, float &lf) {
  const double hd = tau + terr, ld = tau - terr;
  hf = (float)hd;
  if (!((double)hf >= hd)) hf = nextafterf(hf, INFINITY);
  lf = (float)ld;
  if (!((double)lf <= ld)) lf = nextafterf(lf, -INFINITY);
}

// A face's power and height polynomials in FP32 (the FP64 coefficients
// rounded to nearest).  With the coordinates, the coefficients and the at
// most 8 roundings of a chain in FP32, |P32 - P| < 1e-6 Pabs and
// |H32 - H| < 1e-6 Habs (about 5e-7 each); a sign is certified when it
// exceeds 1e-5 of its magnitude bound and the bound exceeds 1e-30 (FP32
// underflow then adds at most ~1e-37, far below the threshold).  Undecided
// points take the FP64 path.
struct TriSphereF {
  float D, DA, g[3], gA[3], n[3], nA[3];
  __device__ void init(const TriSphere &f) {
    D = (float)f.D;
    DA = (float)f.DA;
    g[0] = (float)f.g.x; g[1] = (float)f.g.y; g[2] = (float)f.g.z;
    gA[0] = (float)f.gA.x; gA[1] = (float)f.gA.y; gA[2] = (float)f.gA.z;
    n[0] = (float)f.n.x; n[1] = (float)f.n.y; n[2] = (float)f.n.z;
    nA[0] = (float)f.nA.x; nA[1] = (float)f.nA.y; nA[2] = (float)f.nA.z;
  }
  __device__ void eval(float x, float y, float z, float &P, float &PA, float &H, float &HA) const {
    const float pp = x * x + y * y + z * z;
    const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
    P = D * pp - (x * g[0] + y * g[1] + z * g[2]);
    PA = DA * pp + (ax * gA[0] + ay * gA[1] + az * gA[2]);
    H = x * n[0] + y * n[1] + z * n[2];
    HA = ax * nA[0] + ay * nA[1] + az * nA[2];
  }
};
// Key interval of a point whose FP32 signs are certified: the true P lies
// within 1e-6 PA of P32 and |H| within 1e-6 HA of |H32| (< |H32| / 10), so
// tau = P / |H| lies between the quotients of the extreme values; the
// roundings of the sums, the reciprocals (1 ulp) and the products stay
// inside the final 1e-6 widening.
__device__ inline void key_bounds_32(float Pv, float PA, float H, float HA, float &hf, float &lf) {
  const float ah = fabsf(H), eP = 1e-6f * PA, eH = 1e-6f * HA;
  const float rlo = __builtin_amdgcn_rcpf(ah + eH), rhi = __builtin_amdgcn_rcpf(ah - eH);
  float lo, hi;
  if (Pv > 0.0f) {
    lo = (Pv - eP) * rlo;
    hi = (Pv + eP) * rhi;
  } else {
    lo = (Pv - eP) * rhi;
    hi = (Pv + eP) * rlo;
  }
  lo = lo - 1e-6f * fabsf(lo) - 1e-37f;
  hi = hi + 1e-6f * fabsf(hi) + 1e-37f;
  if (!(fabsf(lo) < INFINITY && fabsf(hi) < INFINITY)) {
    lo = -INFINITY;
    hi = INFINITY;
  }
  hf = hi;
  lf = lo;
}

// Apexes of one side of a face: the outside points ahead of the side whose
// sweep key is among the s_max-3 smallest (fewer when inside points are
// already ahead), each checked against the inside points behind the side.
// One pass over the region fills the side's sweep list; the candidates are
// packed into a mask before any tetrahedron is classified.
// Capsule holding every sphere of a face's pencil that can be critical:
// centres c3 + t n with |t| <= T, radii <= Rb.  A point farther than T + Rb
// from c3 is outside all of them (checked with a relative margin).
struct Capsule {
  V3 c3;
  double bound2;  // (T + Rb)^2 with margin; INFINITY disables the test
  __device__ bool outside(const V3 &p) const {
    const V3 d = sub(p, c3);
    return dot(d, d) > bound2;
  }
};

// The face passes scan the region in chunks of 64 slots: first the capsule
// test of every slot of the chunk (a bit mask), then the point work for the
// set bits only, in ascending slot order, exactly the order of a plain scan.
// The lanes of a wave run the point work as often as the busiest lane of the
// chunk needs it, instead of once for every slot that any lane needs.
// tests counts the slots tested (the face's own vertices excluded).
__device__ inline uint64_t capsule_chunk(const Work &w, const Capsule &cap, uint32_t s0, uint32_t M, uint32_t vi,
                                         uint32_t wi, uint32_t &tests) {
  const uint32_t cn = min(64u, M - s0);
  uint64_t in = 0;
#pragma unroll 4
  for (uint32_t j = 0; j < cn; ++j) {
    const uint32_t gi = s0 + j;
    const bool keep = gi != vi && gi != wi && !cap.outside(gpt(w, gi));
    in |= (uint64_t)keep << j;
  }
  tests += cn - (vi - s0 < cn ? 1u : 0u) - (wi - s0 < cn ? 1u : 0u);
  return in;
}

// The same test on the FP32 coordinates: centre rounded to FP32 and the
// squared bound widened by 1e-5 (rounded up), which covers the FP32
// rounding of the coordinates, the centre and the distance (about 3e-6 of
// the bound for a point near it).
__device__ inline uint64_t capsule_chunk_f(const Work &w, float cx, float cy, float cz, float b2, uint32_t s0,
                                           uint32_t M, uint32_t vi, uint32_t wi, uint32_t &tests32) {
  const uint32_t cn = min(64u, M - s0);
  uint64_t in = 0;
#pragma unroll 4
  for (uint32_t j = 0; j < cn; ++j) {
    const uint32_t gi = s0 + j;
    const float dx = w.fx[gi] - cx, dy = w.fy[gi] - cy, dz = w.fz[gi] - cz;
    const bool keep = gi != vi && gi != wi && !(dx * dx + dy * dy + dz * dz > b2);
    in |= (uint64_t)keep << j;
  }
  tests32 += cn - (vi - s0 < cn ? 1u : 0u) - (wi - s0 < cn ? 1u : 0u);
  return in;
}

// Per-wave work queues of the face phase (LDS).  The first pass of every
// face runs one face per lane; the per-side passes and the tetrahedra it
// leads to are queued and then taken one per lane, so the lanes of a wave
// run passes of equal length instead of waiting for the face with the most
// apexes.  A full queue falls back to doing the work in place.
constexpr uint32_t kSideQ = 128;  // >= 2 items for each of a wave's 64 faces
constexpr uint32_t kTetQ = 256;
struct WaveQueues {
  uint32_t *side;    // vi | wi << 10 | side << 20 | cin << 21 (region slots, < 1024)
  float *side_r;     // the face's rmax (a float, see face_feasible): the side pass rebuilds exactly the face pass's capsule
  uint32_t *tet;     // vi | wi << 10 | zi << 20
  uint32_t *nside, *ntet;
};

__device__ void process_tet(const Params &P, const Work &w, uint32_t u, uint32_t M, uint32_t vi, uint32_t wi,
                            uint32_t zi, uint32_t *region_slot);

__device__ void face_side_apexes(const Params &P, const Work &w, uint32_t u, uint32_t M, const TriSphere &f,
                                 const Capsule &cap, uint32_t vi, uint32_t wi, uint32_t wid, int side,
                                 int cin_side, int K, uint32_t *region_slot, WaveQueues *q) {
  const int need = K - cin_side;  // slots left for points swept before z
  if (need <= 0) return;
  SweepList top;
  BehindList beh;
  top.init();
  beh.init();
  bool undecided = false;
  uint32_t tests = 0, evals = 0;
  // a point outside every feasible sphere of the pencil can neither be an
  // apex nor be swept before one (dropping it only loosens the threshold)
#pragma unroll 1
  for (uint32_t s0 = 0; s0 < M; s0 += 64) {
    uint64_t in = capsule_chunk(w, cap, s0, M, vi, wi, tests);
#pragma unroll 1
    while (in) {
      const uint32_t s = s0 + (uint32_t)(__ffsll((long long)in) - 1);
      in &= in - 1;
      const V3 p = gpt(w, s);
      evals++;
      double Pw, PA, H, HA;
      f.power(p, Pw, PA);
      f.height(p, H, HA);
      const int sp = cert_sign(Pw, PA), sh = cert_sign(H, HA);
      undecided |= sp == 0 || sh == 0;
      const bool ahead = side == 0 ? sh > 0 : sh < 0;
      double tau, terr;
      sweep_key(Pw, PA, H, HA, tau, terr);
      if (sp < 0) {
        // inside the triangle's sphere, behind this side: inside the
        // pencil's spheres until the sweep key reaches -tau
        if (!ahead) beh.insert(-tau - terr);
      } else if (ahead) {
        top.insert(tau, terr, s);
      }
    }
  }
  add_flops(w, (uint64_t)tests * kFlopCapsule + (uint64_t)evals * (kFlopFacePoint + kFlopSweepKey));
  // the face pass certified every sign with the same arithmetic
  if (undecided) return;
  const double thr = top.threshold(need);
  // evicted_lo stays +inf while no point has left or missed the list (every
  // candidate is then listed, also when fewer than `need` are: thr = +inf)
  if (top.evicted_lo != INFINITY && top.evicted_lo <= thr) {
    // a point outside the list may still qualify (near-equal keys): full scan
    tests = evals = 0;
#pragma unroll 1
    for (uint32_t s = 0; s < M; ++s) {
      const uint32_t zi = s;
      if (zi == vi || zi == wi || w.gid[zi] <= wid) continue;
      const V3 p = gpt(w, zi);
      tests++;
      if (cap.outside(p)) continue;
      evals++;
      double Pw, PA, H, HA;
      f.power(p, Pw, PA);
      f.height(p, H, HA);
      if (cert_sign(Pw, PA) <= 0) continue;
      const int sh = cert_sign(H, HA);
      if ((side == 0 && sh <= 0) || (side == 1 && sh >= 0)) continue;
      double tau, terr;
      sweep_key(Pw, PA, H, HA, tau, terr);
      if (tau - terr > thr) continue;
      if (cin_side + beh.count_above(tau + terr) >= K) continue;
      process_tet(P, w, u, M, vi, wi, zi, region_slot);
    }
    add_flops(w, (uint64_t)tests * kFlopCapsule + (uint64_t)evals * (kFlopFacePoint + kFlopSweepKey));
    return;
  }
  uint32_t cand = 0;
#pragma unroll
  for (int j = 0; j < SweepList::kCap; ++j) {
    const double terr = (double)top.err[j], h = top.hi[j];
    // inside points behind: counted while their key exceeds z's
    const bool ok = top.slot[j] != kInvalid && h - 2.0 * terr <= thr && cin_side + beh.count_above(h) < K;
    cand |= ok ? (1u << j) : 0u;
  }
  while (cand) {
    const int j = __ffs(cand) - 1;
    cand &= cand - 1;
    uint32_t slot = 0;  // selected by static index, the list stays in registers
#pragma unroll
    for (int q = 0; q < SweepList::kCap; ++q) slot = q == j ? top.slot[q] : slot;
    const uint32_t zi = slot;
    if (w.gid[zi] <= wid) continue;
    if (q) {
      const uint32_t at = atomicAdd(q->ntet, 1u);
      if (at < kTetQ) {
        q->tet[at] = vi | (wi << 10) | (zi << 20);
        continue;
      }
    }
    process_tet(P, w, u, M, vi, wi, zi, region_slot);
  }
}

// The pencil capsule of face (u, v, w) for critical radii up to rmax.
__device__ inline Capsule face_capsule(const TriSphere &f, double rmax) {
  Capsule cap;
  cap.c3 = {0.0, 0.0, 0.0};
  cap.bound2 = INFINITY;
  if (f.D > 1e-6 * f.DA) {
    const double s2 = 0.5 / f.D;
    cap.c3 = {f.g.x * s2, f.g.y * s2, f.g.z * s2};
    const double r3 = sqrt(dot(cap.c3, cap.c3));
    const double T = sqrt(fmax(0.0, rmax * rmax - r3 * r3));
    const double Rb = fmax(rmax, r3);
    cap.bound2 = (T + Rb) * (T + Rb) * (1.0 + 1e-6);
  }
  return cap;
}

// Apexes of one side from the FP32 lists its face's first pass filled (the
// tail of face_side_apexes).  When a point left or missed the list with a
// lower bound within the threshold, the side takes its own pass instead
// (queued, with the full scan that case needs).
__device__ inline void fused_side(const Params &P, const Work &w, uint32_t u, uint32_t M, uint32_t vi, uint32_t wi,
                                  uint32_t wid, int side, int cin_side, int K, const SweepListF &top,
                                  const BehindListF &beh, double rmax, uint32_t *region_slot, WaveQueues *q) {
  const int need = K - cin_side;
  const float thr = top.threshold(need);
  if (top.evicted_lo != INFINITY && top.evicted_lo <= thr) {
    if (!q) {
      // in place (global-memory path): the face's sphere rebuilt (same
      // arithmetic, same values), then the side's own pass
      TriSphere f;
      f.init(gpt(w, vi), gpt(w, wi));
      const Capsule cap = face_capsule(f, rmax);
      face_side_apexes(P, w, u, M, f, cap, vi, wi, wid, side, cin_side, K, region_slot, nullptr);
      return;
    }
    // at most 2 items for each of the wave's 64 faces: the queue cannot fill
    const uint32_t at = atomicAdd(q->nside, 1u);
    if (at < kSideQ) {
      q->side[at] = vi | (wi << 10) | ((uint32_t)side << 20) | ((uint32_t)cin_side << 21);
      q->side_r[at] = (float)rmax;  // exact: rmax is a float value
    } else {
      atomicAdd(&P.counters[4], 1ull);  // reported by the host as an internal error
    }
    return;
  }
  uint32_t cand = 0;
#pragma unroll
  for (int j = 0; j < SweepListF::kCap; ++j) {
    // inside points behind: counted while their key exceeds z's
    const bool ok = top.ent[j] != kInvalid && entry_lo(top.hi[j], top.ent[j]) <= thr &&
                    cin_side + beh.count_above(top.hi[j]) < K;
    cand |= ok ? (1u << j) : 0u;
  }
  while (cand) {
    const int j = __ffs(cand) - 1;
    cand &= cand - 1;
    uint32_t ent = 0;  // selected by static index, the list stays in registers
#pragma unroll
    for (int qq = 0; qq < SweepListF::kCap; ++qq) ent = qq == j ? top.ent[qq] : ent;
    const uint32_t zi = entry_slot(ent);
    if (w.gid[zi] <= wid) continue;
    if (q) {
      const uint32_t at = atomicAdd(q->ntet, 1u);
      if (at < kTetQ) {
        q->tet[at] = vi | (wi << 10) | (zi << 20);
        continue;
      }
    }
    process_tet(P, w, u, M, vi, wi, zi, region_slot);
  }
}

// Face (u, v, w): its triangle event, then the apexes of each side.  One
// pass certifies every side and insideness, builds the triangle's interior,
// counts the inside points ahead of each side and fills both sides' sweep
// lists, from which the apexes follow (fused_side).  QUEUED (LDS paths):
// tetrahedra and rescans go to the wave's queues; otherwise in place.  F32:
// the pass runs on the region's FP32 copies.
template <bool QUEUED, bool F32>
__device__ void process_face(const Params &P, const Work &w, uint32_t u, uint32_t M, uint32_t vi,
                             uint32_t wi, double rmax, uint32_t *region_slot, WaveQueues *q) {
  // the sphere is built from its points in PointId order, so its FP64 level
  // (and every later tetrahedron's) does not depend on the enumeration order
  uint32_t ids[2] = {w.gid[vi], w.gid[wi]};
  if (ids[0] > ids[1]) { uint32_t t = ids[0]; ids[0] = ids[1]; ids[1] = t; uint32_t q = vi; vi = wi; wi = q; }
  TriSphere f;
  const int wc = f.init(gpt(w, vi), gpt(w, wi));
  add_flops(w, kFlopTriInit);
  if (!f.nondegenerate()) {
    // collinear or undecided: a triangle cannot be a support; an undecided
    // face goes to the host together with its apexes
    if (cert_sign(f.D, f.DA) == 0) emit_fallback(P, w, u, 3, ids, 3, M, region_slot);
    return;
  }
  const uint32_t wid = ids[1];
  const int smax = P.s_max;
  const int K = smax - 3;  // an apex must be among the first K points swept
  // pencil capsule: a critical sphere through u, v, w has its direction in a
  // cell shared by v and w, so its radius is at most rmax; its centre is
  // c3 + t n with r3^2 + t^2 <= rmax^2.  Only well-conditioned faces use it.
  const Capsule cap = face_capsule(f, rmax);
  // the triangle's level, before the pass (the vertices need not stay live)
  const double tri_level = f.level(), tri_level_err = f.level_err();
  int cin0 = 0, cin1 = 0;  // inside the triangle's sphere, ahead of side 0 / 1
  int total_in = 0;
  // region slots of the first 8 inside points, 16 bits each, 2 per word (a
  // triangle event has at most s_max - 3 <= 8 interior points)
  uint32_t tri_pk[4] = {0u, 0u, 0u, 0u};
  uint32_t tests = 0, evals = 0;
  SweepListF top0, top1;  // outside points ahead of side 0 / 1
  BehindListF beh0, beh1;  // inside points behind side 0 / 1
  top0.init();
  top1.init();
  beh0.init();
  beh1.init();
  constexpr uint32_t kFlopPoint = kFlopFacePoint + kFlopSweepKey;
  uint32_t tests32 = 0, evals32 = 0, keys32 = 0;  // FP32 work (F32 path)
  // the FP32 path: capsule and face polynomials in FP32, FP64 only for the
  // points whose FP32 signs are not certified (no FP64 capsule tests counted)
  TriSphereF f32;
  float ccx = 0.0f, ccy = 0.0f, ccz = 0.0f, cb2 = INFINITY;
  if (F32) {
    f32.init(f);
    ccx = (float)cap.c3.x;
    ccy = (float)cap.c3.y;
    ccz = (float)cap.c3.z;
    const double b2 = cap.bound2 * (1.0 + 1e-5);
    cb2 = (float)b2;
    if (!((double)cb2 >= b2)) cb2 = nextafterf(cb2, INFINITY);
  }
  // points outside the capsule are outside the triangle's sphere and every
  // feasible one
#pragma unroll 1
  for (uint32_t s0 = 0; s0 < M; s0 += 64) {
    uint64_t in = F32 ? capsule_chunk_f(w, ccx, ccy, ccz, cb2, s0, M, vi, wi, tests32)
                      : capsule_chunk(w, cap, s0, M, vi, wi, tests);
#pragma unroll 1
    while (in) {
      const uint32_t gi = s0 + (uint32_t)(__ffsll((long long)in) - 1);
      in &= in - 1;
      int sp = 0, sh = 0;
      float hf = 0.0f, lf = 0.0f;
      if (F32) {
        float P32, PA32, H32, HA32;
        f32.eval(w.fx[gi], w.fy[gi], w.fz[gi], P32, PA32, H32, HA32);
        sp = cert32(P32, PA32);
        sh = cert32(H32, HA32);
        evals32++;
        if (sp != 0 && sh != 0) {
          key_bounds_32(P32, PA32, H32, HA32, hf, lf);
          keys32++;
        }
      }
      if (sp == 0 || sh == 0) {
        const V3 p = gpt(w, gi);
        evals++;
        double Pw, PA, H, HA;
        if (F32) {
          // rare: the FP64 sphere is rebuilt (same arithmetic, same values)
          // rather than kept live through the pass
          TriSphere f64;
          f64.init(gpt(w, vi), gpt(w, wi));
          f64.power(p, Pw, PA);
          f64.height(p, H, HA);
        } else {
          f.power(p, Pw, PA);
          f.height(p, H, HA);
        }
        sp = cert_sign(Pw, PA);
        sh = cert_sign(H, HA);
        // every sign the side passes use is certified here, before anything is emitted
        if (sp == 0 || sh == 0) {
          add_flops(w, (uint64_t)tests * kFlopCapsule + (uint64_t)evals * kFlopPoint);
          add_flops32(w, (uint64_t)tests32 * kF32Capsule + (uint64_t)evals32 * kF32FacePoint + (uint64_t)keys32 * kF32Key);
          emit_fallback(P, w, u, 3, ids, 3, M, region_slot);
          return;
        }
        double tau, terr;
        sweep_key(Pw, PA, H, HA, tau, terr);
        key_bounds_f(tau, terr, hf, lf);
      }
      if (sp > 0) {
        // outside: ahead of side 0 when sh > 0, of side 1 otherwise
        if (sh > 0) top0.insert(hf, lf, gi);
        else top1.insert(hf, lf, gi);
      } else {
        // inside: behind side 1 when sh > 0, behind side 0 otherwise; -hf
        // is a lower bound of -tau - terr
        if (sh > 0) beh1.insert(-hf);
        else beh0.insert(-hf);
      }
      if (sp > 0) continue;
      if (total_in < 8) {
        const uint32_t word = (uint32_t)total_in >> 1, shift = ((uint32_t)total_in & 1u) * 16u;
#pragma unroll
        for (int k = 0; k < 4; ++k) tri_pk[k] |= (uint32_t)k == word ? gi << shift : 0u;
      }
      total_in++;
      cin0 += sh > 0;
      cin1 += sh < 0;
      // both sides saturated and the triangle out of rank: nothing left here
      if (cin0 >= K && cin1 >= K && total_in + 3 > smax) {
        add_flops(w, (uint64_t)tests * kFlopCapsule + (uint64_t)evals * kFlopPoint);
        add_flops32(w, (uint64_t)tests32 * kF32Capsule + (uint64_t)evals32 * kF32FacePoint + (uint64_t)keys32 * kF32Key);
        return;
      }
    }
  }
  add_flops(w, (uint64_t)tests * kFlopCapsule + (uint64_t)evals * kFlopPoint);
  add_flops32(w, (uint64_t)tests32 * kF32Capsule + (uint64_t)evals32 * kF32FacePoint + (uint64_t)keys32 * kF32Key);
  // triangle event
  if (wc > 0 && total_in + 3 <= smax) {
    atomicAdd(w.cand, 1u);
    uint32_t tri_int[kMaxInterior];
#pragma unroll
    for (int j = 0; j < kMaxInterior; ++j)
      tri_int[j] = j < 8 && j < total_in ? w.gid[(tri_pk[j / 2] >> ((j % 2) * 16)) & 0xFFFFu] : kInvalid;
    sort_static(tri_int, total_in);
    emit_event(P, u, ids, 3, total_in, tri_int, tri_level, tri_level_err);
  } else if (wc == 0 && total_in + 3 <= smax) {
    emit_fallback(P, w, u, 3, ids, 3, M, region_slot);
    return;
  }
  if (K <= 0) return;
  // queued on the LDS paths, in place on the global-memory path
  WaveQueues *qq = QUEUED ? q : nullptr;
  if (cin0 < K) fused_side(P, w, u, M, vi, wi, wid, 0, cin0, K, top0, beh0, rmax, region_slot, qq);
  if (cin1 < K) fused_side(P, w, u, M, vi, wi, wid, 1, cin1, K, top1, beh1, rmax, region_slot, qq);
}

// A queued side of a face: the face's sphere rebuilt (same arithmetic, same
// values), then the side's pass.
__device__ void side_item(const Params &P, const Work &w, uint32_t u, uint32_t M, uint32_t lo, double rmax,
                          uint32_t *region_slot, WaveQueues *q) {
  const uint32_t vi = lo & 1023u, wi = (lo >> 10) & 1023u, side = (lo >> 20) & 1u, cin_side = lo >> 21;
  TriSphere f;
  f.init(gpt(w, vi), gpt(w, wi));
  add_flops(w, kFlopTriInit);
  const Capsule cap = face_capsule(f, rmax);
  face_side_apexes(P, w, u, M, f, cap, vi, wi, w.gid[wi], (int)side, (int)cin_side, P.s_max - 3, region_slot, q);
}

__device__ void tet_item(const Params &P, const Work &w, uint32_t u, uint32_t M, uint32_t item,
                         uint32_t *region_slot) {
  const uint32_t vi = item & 1023u, wi = (item >> 10) & 1023u, zi = item >> 20;
  process_tet(P, w, u, M, vi, wi, zi, region_slot);
}

// Cell direction and angular radius in FP32, rounded so that the cell only
// grows (cos delta down, sin delta up): every FP32 test below errs towards a
// larger certificate radius and a larger region (the table is rounded on
// the host, see make_cells_f).  Reciprocals and square
// roots use the hardware approximations (1 ulp); their error, like the
// rounding of the products, stays far inside the slack of the tests (1e-5 on
// the mask cosines, 1e-4 on the certificate radius; cancellation is bounded
// by skipping covers with cos(theta + delta) < 0.01).  Lengths below kTinyLen
// (where the approximations flush) take the conservative branch.
constexpr float kTinyLen = 1e-30f;
struct CellF {
  float dx, dy, dz, cosd, sind;
};
__device__ inline CellF cell_f(const float *cf) {
  const float4 a = *reinterpret_cast<const float4 *>(cf);
  return {a.x, a.y, a.z, a.w, cf[4]};
}

// Cells whose union of closed tangent balls (radius rho'_C) contains gathered
// point i: cos of the smallest angle between its direction and the cell.
// FP32 with relative and absolute slack; a point may only be included too
// often, which costs work but never a support.
// rmax2 bounds 2 rho'_C over every cell from above: a point beyond it is in
// no cell's balls; wlim[wd] bounds it over the 64 cells of mask word wd.
__device__ bool member_mask(const Params &P, const Work &w, uint32_t i, float rmax2, const float *wlim,
                            uint64_t mask[6]) {
  const float4 q = w.fc[i];
  for (int k = 0; k < 6; ++k) mask[k] = 0;
  if (!(q.w > kTinyLen)) {  // distance lost to underflow: member everywhere
    for (int k = 0; k < 6; ++k) mask[k] = ~0ull;
    return true;
  }
  if (q.w > rmax2) return false;
  const float inv = __builtin_amdgcn_rcpf(q.w);
  bool any = false;
  // one mask word per 64 cells, each built in a register; a word whose
  // largest radius the point exceeds stays empty
#pragma unroll
  for (int wd = 0; wd < kCells / 64; ++wd) {
    if (q.w > wlim[wd]) continue;
    add_flops32(w, 64ull * kF32Mask);
    uint64_t bits = 0;
#pragma unroll 1
    for (int b = 0; b < 64; ++b) {
      const int c = 64 * wd + b;
      const CellF cf = cell_f(P.cellf + 8 * c);
      const float ct = (q.x * cf.dx + q.y * cf.dy + q.z * cf.dz) * inv;
      float cm;  // cos(max(theta - delta, 0))
      if (ct >= cf.cosd) cm = 1.0f;
      else cm = ct * cf.cosd + __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - ct * ct)) * cf.sind;
      const float rho = w.rho[c];
      if (q.w <= 2.0f * rho * (cm + 1e-5f) * (1.0f + 1e-5f)) bits |= 1ull << b;
    }
    mask[wd] = bits;
    any |= bits != 0;
  }
  return any;
}

constexpr int kCoverCap = 10;  // s_max - 1 <= 10

// Face (u, v, w) of region slots si, sj: false when v and w share no cell or
// when the triangle's circumradius exceeds every shared cell's certificate
// radius (well-conditioned faces only; FP64 r2 is then accurate to ~1e-10).
// rmax, the largest radius over the shared cells, is rounded up to a float
// value, so that the queued side passes can keep it in a float and rebuild
// exactly the first pass's capsule (a larger bound only prunes less).
__device__ inline bool face_feasible(const Work &w, uint32_t si, uint32_t sj, double &rmax, float *key = nullptr) {
  bool share = false;
  for (int k = 0; k < 6; ++k) share |= (w.rmask[6 * si + k] & w.rmask[6 * sj + k]) != 0;
  rmax = 0;
  if (!share) return false;
  const V3 a = gpt(w, si), b = gpt(w, sj);
  const double aa = dot(a, a), bb = dot(b, b), ab = dot(a, b);
  const double D = aa * bb - ab * ab;
  const V3 amb = sub(a, b);
  const double r2 = D > 0 ? aa * bb * dot(amb, amb) / (4.0 * D) : INFINITY;
  for (int k = 0; k < 6; ++k) {
    uint64_t mm = w.rmask[6 * si + k] & w.rmask[6 * sj + k];
    while (mm) {
      const int bit = __ffsll((long long)mm) - 1;
      mm &= mm - 1;
      rmax = fmax(rmax, w.rho[64 * k + bit]);
    }
  }
  float rf = (float)rmax;
  if ((double)rf < rmax) rf = nextafterf(rf, INFINITY);
  rmax = (double)rf;
  const bool well = D > 1e-6 * (aa * bb + fabs(ab) * fabs(ab));
  if (key) {
    // the squared extent of the face's pencil capsule (an estimate, it only
    // orders the faces into waves of similar work)
    const double r3 = sqrt(r2), t = sqrt(fmax(rmax * rmax - r2, 0.0)) + fmax(rmax, r3);
    *key = well ? (float)(t * t) : INFINITY;
  }
  return !(well && r2 > rmax * rmax * (1.0 + 1e-6));
}

// Pairs, faces and tetrahedra of anchor u over its compacted region (M
// points at the front of the point arrays, masks by region slot, cell radii
// in w.rho); *sK == 0 and *sregion == kInvalid on entry, visible to every
// lane.  Marks the anchor done.
template <bool BIG, bool F32>
__device__ void anchor_faces(const Params &P, const Work &w, uint32_t u, uint32_t M, uint32_t *sK_,
                             uint32_t *sregion_) {
  uint32_t &sK = *sK_;
  uint32_t &sregion = *sregion_;
  const int t = threadIdx.x;
  // candidate vertices (ids above the anchor); their order only decides
  // which lane takes which face (a rank sort by direction cell, to give
  // neighbouring lanes overlapping pencil capsules, measured slower); the
  // FP32 copies of the region's coordinates for the face passes
  for (uint32_t s = t; s < M; s += blockDim.x) {
    if (w.gid[s] > u) w.kidx[atomicAdd(&sK, 1u)] = s;
    if (F32) {
      w.fx[s] = (float)w.gx[s];
      w.fy[s] = (float)w.gy[s];
      w.fz[s] = (float)w.gz[s];
    }
  }
  __syncthreads();
  const uint32_t Kc = sK;
  // 3. pairs (u, v), v > u
  for (uint32_t k = t; k < Kc; k += blockDim.x) process_pair(P, w, u, M, w.kidx[k], &sregion);
  // 4. faces (u, v, w), u < v < w, whose members share a cell.  Each wave
  // takes 64 candidate faces at a time: first passes one per lane, then the
  // queued sides and tetrahedra one per lane (LDS path; the global-memory
  // path works in place).
  __shared__ uint32_t q_side[kBlock / 64][kSideQ];
  __shared__ float q_side_r[kBlock / 64][kSideQ];
  __shared__ uint32_t q_face[kBlock / 64][128];
  __shared__ float q_key[kBlock / 64][128];
  __shared__ uint32_t q_tet[kBlock / 64][kTetQ];
  __shared__ uint32_t q_n[kBlock / 64][2];
  const int wave = t >> 6, lane = t & 63, nwaves = (int)(blockDim.x >> 6);
  WaveQueues wq{q_side[wave], q_side_r[wave], q_tet[wave], &q_n[wave][0], &q_n[wave][1]};
  WaveQueues *q = BIG ? nullptr : &wq;
  if (lane == 0) { q_n[wave][0] = 0; q_n[wave][1] = 0; }
  wave_sync();
  const uint64_t F = (uint64_t)Kc * (Kc - 1) / 2;
  // candidate face fidx -> region slots of (v, w): the pair (i < j) of the
  // strictly lower triangle
  auto face_of = [&](uint64_t fidx, uint32_t &si, uint32_t &sj) {
    uint64_t j = (uint64_t)((1.0 + sqrt(1.0 + 8.0 * (double)fidx)) * 0.5);
    while (j * (j - 1) / 2 > fidx) j--;
    while ((j + 1) * j / 2 <= fidx) j++;
    const uint64_t i = fidx - j * (j - 1) / 2;
    si = w.kidx[i];
    sj = w.kidx[j];
  };
  if (BIG) {
    for (uint64_t base = (uint64_t)wave * 64; base < F; base += (uint64_t)nwaves * 64) {
      const uint64_t fidx = base + lane;
      if (fidx >= F) continue;
      uint32_t si, sj;
      face_of(fidx, si, sj);
      double rmax;
      if (face_feasible(w, si, sj, rmax)) process_face<false, false>(P, w, u, M, si, sj, rmax, &sregion, q);
    }
  } else {
    // Each wave first collects up to 128 feasible faces in its LDS list (the
    // faces that fail the shared-cell or radius test cost only the test),
    // orders them by the extent of their pencil capsules, then runs the
    // first passes of the 64 largest one per lane (lanes of similar work),
    // then the queued sides and tetrahedra; the smaller ones wait for the
    // next round.  The lists keep region slots only; a lane recomputes the
    // face's rmax (same arithmetic, same value).
    uint32_t nf = 0;  // faces in the wave's list (uniform)
    uint64_t cbase = (uint64_t)wave * 64;
    const uint64_t cstride = (uint64_t)nwaves * 64;
    for (;;) {
      while (nf <= 64 && cbase < F) {
        const uint64_t fidx = cbase + lane;
        bool ok = false;
        uint32_t si = 0, sj = 0;
        float key = 0.0f;
        if (fidx < F) {
          face_of(fidx, si, sj);
          double rmax;
          ok = face_feasible(w, si, sj, rmax, &key);
        }
        const uint64_t m = __ballot(ok);
        if (ok) {
          const uint32_t at = nf + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
          q_face[wave][at] = si | (sj << 10);
          q_key[wave][at] = key;
        }
        nf += (uint32_t)__popcll(m);
        cbase += cstride;
      }
      if (nf == 0) break;
      wave_sync();
      if (nf > 64) {
        // rank sort of the list by key (ties by position), ascending
        uint32_t fv[2], rk[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const uint32_t i = (uint32_t)lane + 64u * h;
          fv[h] = 0;
          rk[h] = 0;
          if (i < nf) {
            const float ki = q_key[wave][i];
            fv[h] = q_face[wave][i];
            uint32_t r = 0;
            for (uint32_t j = 0; j < nf; ++j) {
              const float kj = q_key[wave][j];
              r += kj < ki || (kj == ki && j < i);
            }
            rk[h] = r;
          }
        }
        wave_sync();
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if ((uint32_t)lane + 64u * h < nf) q_face[wave][rk[h]] = fv[h];
        wave_sync();
      }
      const uint32_t take = min(nf, 64u);
      nf -= take;
      if ((uint32_t)lane < take) {
        const uint32_t fv = q_face[wave][nf + lane];
        const uint32_t si = fv & 1023u, sj = fv >> 10;
        double rmax;
        face_feasible(w, si, sj, rmax);
        process_face<true, F32>(P, w, u, M, si, sj, rmax, &sregion, q);
      }
      wave_sync();
      const uint32_t ns = min(q_n[wave][0], kSideQ);
      for (uint32_t it = lane; it < ns; it += 64)
        side_item(P, w, u, M, q_side[wave][it], (double)q_side_r[wave][it], &sregion, q);
      wave_sync();
      // whole rounds of 64 tetrahedra (the top of the queue); the rest waits
      // for the next round's, so that the lanes stay busy
      const uint32_t nt = min(q_n[wave][1], kTetQ);
      const uint32_t full = nt & ~63u;
      for (uint32_t it = nt - full + lane; it < nt; it += 64) tet_item(P, w, u, M, q_tet[wave][it], &sregion);
      wave_sync();
      if (lane == 0) { q_n[wave][0] = 0; q_n[wave][1] = nt - full; }
      wave_sync();
    }
    const uint32_t nt = min(q_n[wave][1], kTetQ);
    for (uint32_t it = lane; it < nt; it += 64) tet_item(P, w, u, M, q_tet[wave][it], &sregion);
  }
  __syncthreads();
  if (t == 0) {
    if (*w.cand) atomicAdd(&P.counters[3], (unsigned long long)*w.cand);
    if (w.flops && *w.flops) atomicAdd(&P.counters[5], *w.flops);
    if (w.flops32 && *w.flops32) atomicAdd(&P.counters[6], *w.flops32);
    P.status[u] = ST_DONE;
  }
}

// BIG: arrays in global slabs, regions beyond LDS; the face phase works in
// place instead of through the wave queues.
template <bool BIG, bool F32>
__device__ void anchor_work(const Params &P, const Work &w, uint32_t u) {
  __shared__ uint32_t sM, sK, sflag, sregion;
  const int t = threadIdx.x;
  const uint32_t G = P.gcount[u];
  const uint64_t off = P.goff[u];
  const double ux = P.X[u], uy = P.Y[u], uz = P.Z[u];
  const double Rg = P.radius[u];
  __shared__ uint32_t scan_buf[kBlock];
  if (t == 0) { sM = 0; sK = 0; sflag = 0; sregion = kInvalid; }
  if (G <= 4 * blockDim.x) {
    // load in increasing distance from the anchor (rank sort in LDS), so that
    // every scan below meets the points most likely to be inside first
    for (uint32_t i = t; i < G; i += blockDim.x) {
      uint32_t id = P.glist[off + i];
      double dx = P.X[id] - ux, dy = P.Y[id] - uy, dz = P.Z[id] - uz;
      w.gx[i] = dx * dx + dy * dy + dz * dz;
    }
    __syncthreads();
    uint32_t rank[4];
    for (int r = 0; r < 4; ++r) {
      uint32_t i = t + r * blockDim.x;
      rank[r] = 0;
      if (i >= G) continue;
      const double di = w.gx[i];
      uint32_t cnt = 0;
      for (uint32_t j = 0; j < G; ++j) {
        const double dj = w.gx[j];
        cnt += (dj < di) || (dj == di && j < i);
      }
      rank[r] = cnt;
    }
    __syncthreads();
    for (int r = 0; r < 4; ++r) {
      uint32_t i = t + r * blockDim.x;
      if (i >= G) continue;
      uint32_t id = P.glist[off + i];
      w.gid[rank[r]] = id;
      w.gx[rank[r]] = P.X[id] - ux;
      w.gy[rank[r]] = P.Y[id] - uy;
      w.gz[rank[r]] = P.Z[id] - uz;
    }
  } else {
    for (uint32_t i = t; i < G; i += blockDim.x) {
      uint32_t id = P.glist[off + i];
      w.gid[i] = id;
      w.gx[i] = P.X[id] - ux;
      w.gy[i] = P.Y[id] - uy;
      w.gz[i] = P.Z[id] - uz;
    }
  }
  __syncthreads();
  // FP32 copy of the offsets with their lengths (rounded up), for the
  // certificate and the masks
  for (uint32_t i = t; i < G; i += blockDim.x) {
    const double x = w.gx[i], y = w.gy[i], z = w.gz[i];
    w.fc[i] = make_float4((float)x, (float)y, (float)z, (float)sqrt(x * x + y * y + z * z) * (1.0f + 1e-6f));
  }
  __syncthreads();
  // 1. direction-cell certificate (FP32; every rounding errs towards larger radii)
  const int need_cover = P.s_max - 1;
  // the points come in increasing distance when the rank sort ran: a point at
  // distance l covers at radius >= l / 2 (up to the FP32 rounding, inside the
  // factor below), so once l / 2 exceeds the list's last entry no later point
  // can enter it (and dropping points only enlarges the radius anyway)
  const bool by_distance = G <= 4 * blockDim.x;
  uint32_t covers = 0;  // point-cell tests (FP32 count)
  for (int c = t; c < kCells; c += blockDim.x) {
    const double *cd = P.cells + 8 * c;
    const CellF cf = cell_f(P.cellf + 8 * c);
    const double chord = cd[5];
    float best[kCoverCap];
#pragma unroll
    for (int j = 0; j < kCoverCap; ++j) best[j] = INFINITY;
    for (uint32_t i = 0; i < G; ++i) {
      const float4 q = w.fc[i];
      if (!(q.w > kTinyLen)) continue;  // skipping a point only enlarges the radius
      if (by_distance && q.w * 0.49999f > best[kCoverCap - 1]) break;
      covers++;
      const float ct = (q.x * cf.dx + q.y * cf.dy + q.z * cf.dz) * __builtin_amdgcn_rcpf(q.w);
      if (ct <= 0.0f) continue;
      const float st = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
      const float cs = ct * cf.cosd - st * cf.sind;  // cos(theta + delta)
      if (cs < 0.01f) continue;  // ill-conditioned cover: skipped
      const float r = q.w * __builtin_amdgcn_rcpf(2.0f * cs);
      if (!(r < best[kCoverCap - 1])) continue;
#pragma unroll
      for (int j = kCoverCap - 1; j > 0; --j) best[j] = r < best[j - 1] ? best[j - 1] : (r < best[j] ? r : best[j]);
      best[0] = fminf(best[0], r);
    }
    float bk = INFINITY;
#pragma unroll
    for (int j = 0; j < kCoverCap; ++j) bk = j == need_cover - 1 ? best[j] : bk;
    const double rc = (double)bk * (1.0 + 1e-4);
    double dir[3] = {cd[0], cd[1], cd[2]};
    double uu[3] = {ux, uy, uz};
    double e = INFINITY;
    for (int a = 0; a < 3; ++a) {
      double lo = dir[a] - chord, hi = dir[a] + chord;
      if (lo > 0) e = fmin(e, (P.bhi[a] - uu[a]) / lo);
      else if (hi < 0) e = fmin(e, (P.blo[a] - uu[a]) / hi);
    }
    e = fmax(e, 0.0) * (1.0 + 1e-9) + 1e-300;
    // stored in FP32, rounded up (a larger radius only enlarges the region)
    const double rd = fmin(rc, e);
    float rf = (float)rd;
    if (!((double)rf >= rd)) rf = rd <= (double)FLT_MIN ? FLT_MIN : nextafterf(rf, INFINITY);
    w.rho[c] = rf;
  }
  add_flops32(w, (uint64_t)covers * kF32Cover);
  __syncthreads();
  __shared__ float s_rmax2, s_wlim[kCells / 64];
  if (t == 0) {
    double m = 0, wm[kCells / 64] = {};
    for (int c = 0; c < kCells; ++c) {
      m = fmax(m, w.rho[c]);
      wm[c >> 6] = fmax(wm[c >> 6], w.rho[c]);
    }
    // 2 max rho in FP32, rounded up past the slack of the membership test
    // (over all cells and over each mask word's 64)
    s_rmax2 = (float)(2.0 * m) * 1.0001f;
    for (int wd = 0; wd < kCells / 64; ++wd) s_wlim[wd] = (float)(2.0 * wm[wd]) * 1.0001f;
    if (2.0 * m > Rg) {
      P.need[u] = 2.0 * m;
      P.status[u] = ST_WIDEN;
      sflag = 1;
    }
  }
  __syncthreads();
  if (sflag) return;
  const float rmax2 = s_rmax2;
  // 2. region membership masks, compacted in gathered order (chunks of four
  // points per lane; the LDS path has a single chunk, so every FP32 offset is
  // read before the first mask overwrites the area they share)
  uint32_t members = 0;
  for (uint32_t c0 = 0; c0 < G; c0 += 4 * blockDim.x) {
    const uint32_t g = min(G - c0, 4 * blockDim.x);
    const uint32_t chunk = (g + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = c0 + t * chunk;
    uint64_t mk[4][6];  // static indices only: the masks stay in registers
    bool mem[4];
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) {
      mem[r] = false;
      if (r < chunk && t * chunk + r < g) {
        mem[r] = member_mask(P, w, b0 + r, rmax2, s_wlim, mk[r]);
        cnt += mem[r];
      }
    }
    scan_buf[t] = cnt;
    __syncthreads();
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
      uint32_t v = t >= o ? scan_buf[t - o] : 0;
      __syncthreads();
      scan_buf[t] += v;
      __syncthreads();
    }
    uint32_t slot = members + scan_buf[t] - cnt;
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) {
      if (!mem[r]) continue;
      if (slot < P.rcap) {
        w.ridx[slot] = b0 + r;
#pragma unroll
        for (int k = 0; k < 6; ++k) w.rmask[6 * slot + k] = mk[r][k];
      }
      slot++;
    }
    members += scan_buf[blockDim.x - 1];
    __syncthreads();
  }
  if (t == 0) sM = members;
  __syncthreads();
  const uint32_t M = sM;
  if (M > P.rcap) {
    if (t == 0) P.status[u] = ST_BIG;
    return;
  }
  if (P.big && M > P.region_limit) {
    if (t == 0) {
      P.need[u] = (double)M;  // the region's size, reported with the conditional status
      P.status[u] = ST_HUGE;
    }
    return;
  }
  // the region's points move to the front of the point arrays, in region
  // order: from here on a region slot indexes them directly.  ridx ascends,
  // so every source lies at or after its destination and after the
  // destinations of earlier chunks; each chunk reads before it writes.
  for (uint32_t c0 = 0; c0 < M; c0 += blockDim.x) {
    const uint32_t s = c0 + t;
    double x = 0, y = 0, z = 0;
    uint32_t id = 0;
    if (s < M) {
      const uint32_t gi = w.ridx[s];
      x = w.gx[gi];
      y = w.gy[gi];
      z = w.gz[gi];
      id = w.gid[gi];
    }
    __syncthreads();
    if (s < M) {
      w.gx[s] = x;
      w.gy[s] = y;
      w.gz[s] = z;
      w.gid[s] = id;
    }
    __syncthreads();
  }
  if (P.probe) {
    // the region is known and compacted in the slab: its size is reported,
    // and a region that fits LDS is enumerated there (k_anchor_region)
    if (t == 0) {
      P.need[u] = (double)M;
      P.status[u] = ST_PROBED;
    }
    return;
  }
  anchor_faces<BIG, F32>(P, w, u, M, &sK, &sregion);
}

// The LDS path in two sizes: most anchors fit the small one, which leaves
// room for four waves per SIMD; the wide one (one workgroup per CU) takes the
// anchors whose gather or region outgrow it, before the global-memory path.
template <int GCAP, int RCAP>
__device__ void anchor_lds(const Params &P, uint32_t a) {
  __shared__ double gx[GCAP], gy[GCAP], gz[GCAP];
  __shared__ uint32_t gid[GCAP];
  __shared__ float rho[kCells];
  __shared__ uint32_t ridx[RCAP];
  // the FP32 offsets (16 B per gathered point) live in the mask area until
  // the masks are written
  __shared__ __align__(16) uint64_t rmask[6 * RCAP];
  static_assert(sizeof(uint64_t) * 6 * RCAP >= sizeof(float4) * GCAP, "fc alias");
  __shared__ uint32_t kidx[RCAP];
  if (a >= P.anchor_count) return;
  const uint32_t u = P.anchors[a];
  if (P.gcount[u] > P.gcap) {
    if (threadIdx.x == 0) P.status[u] = ST_BIG;
    return;
  }
  __shared__ uint32_t cand;
  __shared__ unsigned long long flops, flops32;
  if (threadIdx.x == 0) { cand = 0; flops = 0; flops32 = 0; }
  __syncthreads();
  // FP32 region coordinates in the tails of the gather arrays, free once the
  // region is compacted to the front
  static_assert(2 * (GCAP - RCAP) >= RCAP, "FP32 region copies fit the gather tails");
  Work w{gx, gy, gz, gid, rho, ridx, rmask, kidx, reinterpret_cast<float4 *>(rmask), &cand,
         P.count_flops ? &flops : nullptr, reinterpret_cast<float *>(gx + RCAP),
         reinterpret_cast<float *>(gy + RCAP), reinterpret_cast<float *>(gz + RCAP),
         P.count_flops ? &flops32 : nullptr};
  anchor_work<false, true>(P, w, u);
}

__global__ void __launch_bounds__(kBlock, 3) k_anchor_lds(Params P) {
  anchor_lds<kGatherCap, kRegionCap>(P, blockIdx.x);
}
__global__ void __launch_bounds__(kBlock, 2) k_anchor_lds_wide(Params P) {
  anchor_lds<kGatherCapWide, kRegionCapWide>(P, blockIdx.x);
}

// One persistent launch over the first anchor pass of several clouds (a
// batch): the workgroups take anchors from a shared counter until every
// cloud's list is done.  Item i belongs to the job j with
// offsets[j] <= i < offsets[j + 1]; each job carries its cloud's parameters
// (coordinates, gathers, record queues).  The item index is uniform across
// the workgroup, so a job's parameters are read with scalar loads.
__global__ void __launch_bounds__(kBlock, 3) k_anchor_persist(const Params *__restrict__ jobs,
                                                              const uint32_t *__restrict__ offsets, int njobs,
                                                              unsigned int *head) {
  __shared__ uint32_t s_item;
  const uint32_t total = offsets[njobs];
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(head, 1u);
    __syncthreads();
    const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (item >= total) return;
    int lo = 0, hi = njobs - 1;  // last job whose offset is <= item
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (offsets[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const int j = __builtin_amdgcn_readfirstlane(lo);
    anchor_lds<kGatherCap, kRegionCap>(jobs[j], item - offsets[j]);
    __syncthreads();
  }
}

// Offset (in doubles, even) of the FP32 offsets in a global slab of gather
// capacity G; the slab holds big_fc_offset(G) + 2G doubles.
__host__ __device__ inline uint64_t big_fc_offset(uint64_t G) {
  return (3 * G + kCells + 6 * G + (3 * G + 1) / 2 + 1 + 1) & ~1ull;
}

// Anchors whose gather or region exceeds LDS: the same work on global slabs.
__global__ void __launch_bounds__(kBlock) k_anchor_big(Params P) {
  const uint32_t a = blockIdx.x;
  if (a >= P.anchor_count) return;
  const uint32_t u = P.anchors[a];
  double *slab = P.scratch + (uint64_t)a * P.slab_doubles;
  const uint32_t G = P.gcap;
  Work w{};
  w.gx = slab;
  w.gy = slab + G;
  w.gz = slab + 2 * G;
  w.rho = reinterpret_cast<float *>(slab + 3 * G);
  w.rmask = (uint64_t *)(slab + 3 * G + kCells);
  w.gid = (uint32_t *)(slab + 3 * G + kCells + 6 * G);
  w.ridx = w.gid + G;
  w.kidx = w.ridx + G;
  w.fc = reinterpret_cast<float4 *>(slab + big_fc_offset(G));
  // the FP32 region copies take the FP32 offsets' area once the region is
  // compacted (4G floats hold 3G)
  w.fx = reinterpret_cast<float *>(w.fc);
  w.fy = w.fx + G;
  w.fz = w.fy + G;
  __shared__ uint32_t cand;
  __shared__ unsigned long long flops;
  if (threadIdx.x == 0) { cand = 0; flops = 0; }
  __syncthreads();
  w.cand = &cand;
  w.flops = P.count_flops ? &flops : nullptr;
  anchor_work<true, true>(P, w, u);
}

// Regions measured by the global-memory path's probe that fit LDS: the
// anchor's compacted region (coordinates, PointIds, masks) and cell radii are
// loaded from its slab into LDS, and the face phase runs there with the wave
// queues.  Two sizes: regions of up to 640 points (two workgroups per CU) and
// of 641 to 1 023 points (one workgroup per CU, 93 KB of LDS; region slots
// stay below 1 024, as the packed queue items need).  Same parameters and
// slab layout as k_anchor_big's probe launch; anchors not probed, or with
// larger regions, are left to it.
constexpr int kRegionCapHuge = 1023;
template <int RCAP, int RMIN>
__global__ void __launch_bounds__(kBlock, RCAP <= kRegionCapWide ? 2 : 1) k_anchor_region(Params P) {
  const uint32_t a = blockIdx.x;
  if (a >= P.anchor_count) return;
  const uint32_t u = P.anchors[a];
  if (P.status[u] != ST_PROBED || !(P.need[u] <= (double)RCAP) || !(P.need[u] > (double)RMIN)) return;
  const uint32_t M = (uint32_t)P.need[u];
  const uint64_t G = P.gcap;
  const double *slab = P.scratch + (uint64_t)a * P.slab_doubles;
  const uint64_t *smask = (const uint64_t *)(slab + 3 * G + kCells);
  const uint32_t *sgid = (const uint32_t *)(slab + 3 * G + kCells + 6 * G);
  __shared__ double gx[RCAP], gy[RCAP], gz[RCAP];
  __shared__ uint32_t gid[RCAP];
  __shared__ float rho[kCells];
  __shared__ uint64_t rmask[6 * RCAP];
  __shared__ uint32_t kidx[RCAP];
  __shared__ float fx[RCAP], fy[RCAP], fz[RCAP];
  __shared__ uint32_t cand, sK, sregion;
  __shared__ unsigned long long flops;
  const int t = threadIdx.x;
  for (uint32_t i = t; i < M; i += blockDim.x) {
    gx[i] = slab[i];
    gy[i] = slab[G + i];
    gz[i] = slab[2 * G + i];
    gid[i] = sgid[i];
  }
  for (uint32_t i = t; i < 6 * M; i += blockDim.x) rmask[i] = smask[i];
  for (int c = t; c < kCells; c += blockDim.x) rho[c] = reinterpret_cast<const float *>(slab + 3 * G)[c];
  if (t == 0) { cand = 0; sK = 0; sregion = kInvalid; flops = 0; }
  __syncthreads();
  Work w{gx, gy, gz, gid, rho, kidx, rmask, kidx, nullptr, &cand, P.count_flops ? &flops : nullptr, fx, fy, fz};
  anchor_faces<false, true>(P, w, u, M, &sK, &sregion);
}

// Ball count / emit around every anchor of the list.
// Pruning runs in FP32 on the two-level LBVH copy (boxes rounded outward, the
// radius inflated by the rounding of the query and of the FP32 arithmetic; a
// non-finite bound prunes nothing); the leaf test is the FP64 one.
template <bool EMIT, int B>
__global__ void __launch_bounds__(B) k_gather(int64_t n, const WideNode *__restrict__ wide,
                                                const double *__restrict__ LX,
                                                const double *__restrict__ LY,
                                                const double *__restrict__ LZ,
                                                const uint32_t *__restrict__ leaf_id,
                                                const uint32_t *__restrict__ leaf_pos,
                                                const uint32_t *__restrict__ anchors, uint32_t count,
                                                const double *__restrict__ radius, uint32_t *gcount,
                                                const uint64_t *__restrict__ goff, uint32_t *glist,
                                                uint32_t *overflow) {
  // depth-first over 4-way nodes: at most 3 pending siblings per level
  constexpr int kStack = 96;
  __shared__ uint32_t stk[kStack * B];
  const int lane = threadIdx.x;
  uint32_t a = blockIdx.x * blockDim.x + lane;
  if (a >= count) return;
  const uint32_t u = anchors[a];
  const uint32_t p = leaf_pos[u];
  const double q[3] = {LX[p], LY[p], LZ[p]};
  const double R = radius[u];
  const double R2 = R * R * (1.0 + 1e-9);
  const float qf[3] = {(float)q[0], (float)q[1], (float)q[2]};
  const double qmax = fmax(fmax(fabs(q[0]), fabs(q[1])), fabs(q[2]));
  const double rb = sqrt(R2) * (1.0 + 1e-6) + qmax * 1e-6 + 1e-30;
  float Tf = (float)(rb * rb * (1.0 + 1e-5));
  if (!(Tf < 3.0e38f)) Tf = INFINITY;
  uint32_t cnt = 0;
  uint64_t base = EMIT ? goff[u] : 0;
  int sp = 0;
  if (n > 1) stk[(sp++) * B + lane] = 0;
  while (sp) {
    const WideNode &N = wide[stk[(--sp) * B + lane]];
    for (int c = 0; c < 4; ++c) {
      const uint32_t x = N.child[c];
      if (x == kInvalid) continue;
      float d2 = 0.f;
      for (int ax = 0; ax < 3; ++ax) {
        const float d = fmaxf(fmaxf(N.lo[c][ax] - qf[ax], qf[ax] - N.hi[c][ax]), 0.f);
        d2 += d * d;
      }
      if (d2 > Tf) continue;
      if (x & kLeafBit) {
        const uint32_t lp = x & ~kLeafBit;
        if (lp == p) continue;
        const double dx = LX[lp] - q[0], dy = LY[lp] - q[1], dz = LZ[lp] - q[2];
        if (dx * dx + dy * dy + dz * dz <= R2) {
          if (EMIT) glist[base + cnt] = leaf_id[lp];
          cnt++;
        }
      } else if (sp < kStack) {
        stk[(sp++) * B + lane] = x;
      } else {
        atomicOr(overflow, 1u);
      }
    }
  }
  if (!EMIT) gcount[u] = cnt;
}

// The same gather with one wave per anchor, for the widened passes, whose
// radii hold hundreds to thousands of points: the wave pops up to 64 pending
// nodes at once (one per lane), tests their four children, and appends hits
// and pending children with ballot prefix counts, so one anchor's traversal
// is spread over 64 lanes instead of walked by one.  The traversal order is
// fixed, so the count and emit passes visit the same points.
constexpr int kWaveGatherStack = 4096;
constexpr int kWaveGatherBlock = 128;  // two waves, two anchors per workgroup
template <bool EMIT>
__global__ void __launch_bounds__(kWaveGatherBlock) k_gather_wave(int64_t n, const WideNode *__restrict__ wide,
                                                                  const double *__restrict__ LX,
                                                                  const double *__restrict__ LY,
                                                                  const double *__restrict__ LZ,
                                                                  const uint32_t *__restrict__ leaf_id,
                                                                  const uint32_t *__restrict__ leaf_pos,
                                                                  const uint32_t *__restrict__ anchors, uint32_t count,
                                                                  const double *__restrict__ radius, uint32_t *gcount,
                                                                  const uint64_t *__restrict__ goff, uint32_t *glist,
                                                                  uint32_t *overflow) {
  __shared__ uint32_t stk[kWaveGatherBlock / 64][kWaveGatherStack];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t a = blockIdx.x * (kWaveGatherBlock / 64) + wave;
  if (a >= count) return;  // uniform over the wave
  uint32_t *S = stk[wave];
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t u = anchors[a];
  const uint32_t p = leaf_pos[u];
  const double q[3] = {LX[p], LY[p], LZ[p]};
  const double R = radius[u];
  const double R2 = R * R * (1.0 + 1e-9);
  const float qf[3] = {(float)q[0], (float)q[1], (float)q[2]};
  const double qmax = fmax(fmax(fabs(q[0]), fabs(q[1])), fabs(q[2]));
  const double rb = sqrt(R2) * (1.0 + 1e-6) + qmax * 1e-6 + 1e-30;
  float Tf = (float)(rb * rb * (1.0 + 1e-5));
  if (!(Tf < 3.0e38f)) Tf = INFINITY;
  uint32_t cnt = 0, sp = 0;
  const uint64_t base = EMIT ? goff[u] : 0;
  if (n > 1) {
    if (lane == 0) S[0] = 0;
    sp = 1;
  }
  bool lost = false;
  wave_sync();
  while (sp > 0) {
    const uint32_t take = min(sp, 64u);
    sp -= take;
    const uint32_t node = (uint32_t)lane < take ? S[sp + lane] : kInvalid;
    wave_sync();  // every pop is read before a push overwrites it
    const WideNode *N = node != kInvalid ? &wide[node] : nullptr;
    for (int c = 0; c < 4; ++c) {
      bool hit = false, push = false;
      uint32_t x = kInvalid;
      if (N) {
        x = N->child[c];
        if (x != kInvalid) {
          float d2 = 0.f;
          for (int ax = 0; ax < 3; ++ax) {
            const float d = fmaxf(fmaxf(N->lo[c][ax] - qf[ax], qf[ax] - N->hi[c][ax]), 0.f);
            d2 += d * d;
          }
          if (d2 <= Tf) {
            if (x & kLeafBit) {
              const uint32_t lp = x & ~kLeafBit;
              if (lp != p) {
                const double dx = LX[lp] - q[0], dy = LY[lp] - q[1], dz = LZ[lp] - q[2];
                hit = dx * dx + dy * dy + dz * dz <= R2;
                if (hit) x = lp;
              }
            } else {
              push = true;
            }
          }
        }
      }
      const uint64_t mh = __ballot(hit);
      if (EMIT && hit) glist[base + cnt + (uint32_t)__popcll(mh & below)] = leaf_id[x];
      cnt += (uint32_t)__popcll(mh);
      const uint64_t mp = __ballot(push);
      const uint32_t slot = sp + (uint32_t)__popcll(mp & below);
      if (push) {
        if (slot < (uint32_t)kWaveGatherStack) S[slot] = x;
        else lost = true;
      }
      sp = min(sp + (uint32_t)__popcll(mp), (uint32_t)kWaveGatherStack);
    }
    wave_sync();
  }
  if (lost) atomicOr(overflow, 1u);
  if (!EMIT && lane == 0) gcount[u] = cnt;
}

// Initial gather radius of every point: factor x the distance to its kq-th
// nearest neighbour (kq <= 32), found in FP32 over the two-level LBVH copy
// with a register-resident sorted list.  The radius only sizes the first
// gather; the certificate widens any anchor it does not cover.
constexpr int kRadiusK = 32;
constexpr int kRadiusStack = 64;
template <int B>
__global__ void __launch_bounds__(B) k_knn_radius(int64_t n, int kq, float factor, const WideNode *__restrict__ wide,
                                                    const double *__restrict__ LX, const double *__restrict__ LY,
                                                    const double *__restrict__ LZ,
                                                    const uint32_t *__restrict__ leaf_id, double *radius) {
  __shared__ uint32_t stk[kRadiusStack * B];
  const int lane = threadIdx.x;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + lane;
  if (p >= n) return;
  const float q[3] = {(float)LX[p], (float)LY[p], (float)LZ[p]};
  float best[kRadiusK];
#pragma unroll
  for (int j = 0; j < kRadiusK; ++j) best[j] = INFINITY;
  // starting bound: the farthest of kq Morton-order neighbours, a set of kq
  // other points, so the kq-th nearest distance cannot exceed it (inflated
  // past the rounding of the outward FP32 leaf boxes, so the traversal still
  // meets those kq points below the bound)
  float bound = INFINITY;
  if (n > kq) {
    int64_t lo = p - kq / 2;
    if (lo < 0) lo = 0;
    if (lo + kq >= n) lo = n - 1 - kq;
    float m = 0.f;
    for (int64_t j = lo; j <= lo + kq; ++j) {
      if (j == p) continue;
      const float dx = (float)LX[j] - q[0], dy = (float)LY[j] - q[1], dz = (float)LZ[j] - q[2];
      m = fmaxf(m, dx * dx + dy * dy + dz * dz);
    }
    bound = m * (1.0f + 1e-4f) + 1e-30f;
  }
  float cut = bound;  // min(bound, the kq-th smallest so far)
  int sp = 0;
  if (n > 1) stk[(sp++) * B + lane] = 0;
  while (sp) {
    const WideNode &N = wide[stk[(--sp) * B + lane]];
    // leaves are taken at once; inner children are pushed far to near, so
    // the nearest is expanded first and the bound shrinks early
    float cd[4];
    uint32_t cx[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t x = N.child[c];
      cd[c] = -1.f;
      cx[c] = x;
      if (x == kInvalid) continue;
      float d2 = 0.f;
      for (int a = 0; a < 3; ++a) {
        const float d = fmaxf(fmaxf(N.lo[c][a] - q[a], q[a] - N.hi[c][a]), 0.f);
        d2 += d * d;
      }
      if (!(d2 < cut)) continue;
      if (x & kLeafBit) {
        if ((x & ~kLeafBit) == (uint32_t)p) continue;
        // sorted insertion over static indices (the list stays in registers)
#pragma unroll
        for (int j = kRadiusK - 1; j > 0; --j) best[j] = d2 < best[j - 1] ? best[j - 1] : (d2 < best[j] ? d2 : best[j]);
        best[0] = fminf(best[0], d2);
        float c2 = INFINITY;
#pragma unroll
        for (int j = 0; j < kRadiusK; ++j) c2 = j == kq - 1 ? best[j] : c2;
        cut = fminf(bound, c2);
      } else {
        cd[c] = d2;
      }
    }
    // descending distance (a five-comparator network), then push
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const int i = e == 0 ? 0 : e == 1 ? 2 : e == 2 ? 0 : e == 3 ? 1 : 1;
      const int j = e == 0 ? 1 : e == 1 ? 3 : e == 2 ? 2 : e == 3 ? 3 : 2;
      if (cd[i] < cd[j]) {
        const float td = cd[i];
        cd[i] = cd[j];
        cd[j] = td;
        const uint32_t tx = cx[i];
        cx[i] = cx[j];
        cx[j] = tx;
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (!(cd[c] >= 0.f) || !(cd[c] < cut)) continue;
      if (sp < kRadiusStack) stk[(sp++) * B + lane] = cx[c];
    }
  }
  // a subtree dropped by a full stack can only raise the kq-th distance found
  const double d = sqrt((double)cut) * (1.0 + 1e-6);
  radius[leaf_id[p]] = (double)factor * d;
}


// exclusive scan of gcount over the anchor list into goff (single workgroup)
__global__ void k_scan_offsets(const uint32_t *__restrict__ anchors, uint32_t count,
                               const uint32_t *__restrict__ gcount, uint64_t *goff, uint64_t *total) {
  __shared__ uint64_t part[1024];
  const int t = threadIdx.x;
  const uint32_t per = (count + blockDim.x - 1) / blockDim.x;
  const uint32_t b = t * per, e = min(b + per, count);
  uint64_t s = 0;
  for (uint32_t i = b; i < e; ++i) s += gcount[anchors[i]];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < (int)blockDim.x; o <<= 1) {
    uint64_t v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint64_t run = part[t] - s;
  for (uint32_t i = b; i < e; ++i) {
    goff[anchors[i]] = run;
    run += gcount[anchors[i]];
  }
  if (t == (int)blockDim.x - 1) *total = part[t];
}

}  // namespace

// Direction cells: centre, cos/sin of the angular radius, chord bound.
static std::vector<double> make_cells() {
  std::vector<double> out(8 * kCells);
  const int G = 8;
  int q = 0;
  for (int f = 0; f < 6; ++f)
    for (int i = 0; i < G; ++i)
      for (int j = 0; j < G; ++j) {
        auto dir = [&](double s, double t, double v[3]) {
          double a = 2 * s - 1, b = 2 * t - 1, w[3];
          int ax = f / 2;
          double sg = (f % 2) ? -1 : 1;
          w[ax] = sg;
          w[(ax + 1) % 3] = a;
          w[(ax + 2) % 3] = b;
          double nn = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
          for (int k = 0; k < 3; ++k) v[k] = w[k] / nn;
        };
        double c[3];
        dir((i + 0.5) / G, (j + 0.5) / G, c);
        double delta = 0;
        for (int ci = 0; ci < 2; ++ci)
          for (int cj = 0; cj < 2; ++cj) {
            double v[3];
            dir((double)(i + ci) / G, (double)(j + cj) / G, v);
            double dt = c[0] * v[0] + c[1] * v[1] + c[2] * v[2];
            delta = std::max(delta, std::acos(std::min(1.0, dt)));
          }
        delta = delta * 1.01 + 1e-9;
        double *o = out.data() + 8 * q++;
        o[0] = c[0]; o[1] = c[1]; o[2] = c[2];
        o[3] = std::cos(delta);
        o[4] = std::sin(delta);
        o[5] = 2.0 * std::sin(delta / 2.0) * 1.001;
        o[6] = delta;
        o[7] = 0;
      }
  return out;
}

// The FP32 cells of the certificate and the membership masks (8 floats per
// cell: direction xyz, cos delta rounded down, sin delta rounded up, padding),
// so that every FP32 test errs towards a larger cell.
static std::vector<float> make_cells_f(const std::vector<double> &cells) {
  std::vector<float> out(8 * kCells, 0.0f);
  for (int c = 0; c < kCells; ++c) {
    const double *cd = cells.data() + 8 * c;
    float *o = out.data() + 8 * c;
    o[0] = (float)cd[0];
    o[1] = (float)cd[1];
    o[2] = (float)cd[2];
    o[3] = (float)cd[3] * (1.0f - 1e-6f);
    o[4] = (float)cd[4] * (1.0f + 1e-6f) + 1e-7f;
  }
  return out;
}

void supports_device(Context &c, int k_max, std::vector<EventRec> &events,
                     std::vector<FallbackRec> &fbs, std::vector<uint32_t> &fb_region,
                     mhgp_catalog_info *info);

// Workgroup tile (queries per workgroup, one query per lane) of the two
// LBVH traversal kernels of the catalogue, the 32nd-neighbour radius and the
// gather.  The tile is picked per device and cloud-size class (powers of two
// of n) by timing every candidate on the first cloud of the class, once per
// process; both kernels are idempotent, so the timing runs are real work.
constexpr int kTiles[3] = {64, 128, 256};
struct TileChoice {
  int radius = 128, gather = 128;
};
std::mutex g_tile_mu;
std::map<std::pair<int, int>, TileChoice> g_tiles;

template <int B>
void launch_radius(hipStream_t s, Context &c, int kq, CatalogBuffers &Bf) {
  hipLaunchKernelGGL(k_knn_radius<B>, dim3(blocks_for(c.n, B)), dim3(B), 0, s, c.n, kq, 2.2f, c.wide.p, c.LX.p,
                     c.LY.p, c.LZ.p, c.leaf_id.p, Bf.radius.p);
}
void launch_radius(int B, hipStream_t s, Context &c, int kq, CatalogBuffers &Bf) {
  if (B == 64) launch_radius<64>(s, c, kq, Bf);
  else if (B == 256) launch_radius<256>(s, c, kq, Bf);
  else launch_radius<128>(s, c, kq, Bf);
}

template <bool EMIT, int B>
void launch_gather(hipStream_t s, Context &c, CatalogBuffers &Bf, uint32_t cnt, uint32_t *overflow) {
  hipLaunchKernelGGL((k_gather<EMIT, B>), dim3(blocks_for(cnt, B)), dim3(B), 0, s, c.n, c.wide.p, c.LX.p, c.LY.p,
                     c.LZ.p, c.leaf_id.p, c.leaf_pos.p, Bf.anchors.p, cnt, Bf.radius.p, Bf.gcount.p, Bf.goff.p,
                     Bf.glist.p, overflow);
}
template <bool EMIT>
void launch_gather(int B, hipStream_t s, Context &c, CatalogBuffers &Bf, uint32_t cnt, uint32_t *overflow) {
  if (B == 64) launch_gather<EMIT, 64>(s, c, Bf, cnt, overflow);
  else if (B == 256) launch_gather<EMIT, 256>(s, c, Bf, cnt, overflow);
  else launch_gather<EMIT, 128>(s, c, Bf, cnt, overflow);
}

template <bool EMIT>
void launch_gather_wave(hipStream_t s, Context &c, CatalogBuffers &Bf, uint32_t cnt, uint32_t *overflow) {
  const uint32_t per = kWaveGatherBlock / 64;
  hipLaunchKernelGGL((k_gather_wave<EMIT>), dim3((cnt + per - 1) / per), dim3(kWaveGatherBlock), 0, s, c.n, c.wide.p,
                     c.LX.p, c.LY.p, c.LZ.p, c.leaf_id.p, c.leaf_pos.p, Bf.anchors.p, cnt, Bf.radius.p, Bf.gcount.p,
                     Bf.goff.p, Bf.glist.p, overflow);
}

// Times `run(tile)` for every candidate tile; returns the fastest.
template <class F>
int fastest_tile(hipStream_t s, F &&run) {
  hipEvent_t a, b;
  MHGP_HIP(hipEventCreate(&a));
  MHGP_HIP(hipEventCreate(&b));
  int best = 128;
  float best_ms = INFINITY;
  for (int B : kTiles) {
    run(B);  // warm (code objects, caches)
    MHGP_HIP(hipEventRecord(a, s));
    run(B);
    MHGP_HIP(hipEventRecord(b, s));
    MHGP_HIP(hipEventSynchronize(b));
    float ms = 0;
    MHGP_HIP(hipEventElapsedTime(&ms, a, b));
    if (ms < best_ms) { best_ms = ms; best = B; }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

TileChoice tiles_for(Context &c, CatalogBuffers &Bf, int kq, uint32_t *overflow) {
  const std::pair<int, int> key{c.device, (int)std::ilogb((double)std::max<int64_t>(c.n, 1))};
  {
    std::lock_guard<std::mutex> g(g_tile_mu);
    auto it = g_tiles.find(key);
    if (it != g_tiles.end()) return it->second;
  }
  hipStream_t s = c.stream;
  TileChoice t;
  t.radius = fastest_tile(s, [&](int B) { launch_radius(B, s, c, kq, Bf); });
  // the gather's count pass over every point in Morton order, as the first
  // anchor pass takes them
  const uint32_t cnt = (uint32_t)c.n;
  MHGP_HIP(hipMemcpyAsync(Bf.anchors.p, c.leaf_id.p, cnt * 4, hipMemcpyDeviceToDevice, s));
  t.gather = fastest_tile(s, [&](int B) { launch_gather<false>(B, s, c, Bf, cnt, overflow); });
  std::lock_guard<std::mutex> g(g_tile_mu);
  return g_tiles.emplace(key, t).first->second;
}

// Rendezvous of the clouds of one batch: each cloud's thread reaches its
// first anchor pass, registers its parameters and waits; the last of them
// queues one persistent launch over every registered cloud on the batch's
// stream (after each cloud's own stream has reached that point), and every
// cloud's stream then waits for that launch.  A cloud that will not reach
// the pass (an error, a cloud too small for supports) leaves instead.
struct AnchorBatch {
  int device = 0;
  std::mutex mu;
  std::condition_variable cv;
  int expected = 0, arrived = 0;
  bool launched = false;
  std::string error;
  std::vector<char> done;  // per slot: arrived or left
  std::vector<Params> jobs;
  std::vector<uint32_t> counts, off;  // host sources of the launch's copies, kept until the batch ends
  std::vector<hipEvent_t> ready;
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  DevBuf<Params> d_jobs;
  DevBuf<uint32_t> d_off;
  DevBuf<unsigned int> d_head;
  int64_t anchors = 0;
  ~AnchorBatch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipEvent_t e : ready) (void)hipEventDestroy(e);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void launch_locked() {
    try {
      MHGP_HIP(hipSetDevice(device));
      if (!stream) {
        MHGP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        MHGP_HIP(hipEventCreate(&t0));
        MHGP_HIP(hipEventCreate(&t1));
      }
      for (hipEvent_t e : ready) MHGP_HIP(hipStreamWaitEvent(stream, e, 0));
      const int J = (int)jobs.size();
      off.assign(J + 1, 0);
      for (int j = 0; j < J; ++j) off[j + 1] = off[j] + counts[j];
      anchors = off[J];
      d_jobs.ensure(J);
      d_off.ensure(J + 1);
      d_head.ensure(1);
      MHGP_HIP(hipMemcpyAsync(d_jobs.p, jobs.data(), J * sizeof(Params), hipMemcpyHostToDevice, stream));
      MHGP_HIP(hipMemcpyAsync(d_off.p, off.data(), (J + 1) * 4, hipMemcpyHostToDevice, stream));
      MHGP_HIP(hipMemsetAsync(d_head.p, 0, 4, stream));
      int cus = 0;
      MHGP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
      // three workgroups per CU fit the kernel's LDS and registers
      const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(off[J], (uint32_t)std::max(cus, 1) * 3));
      MHGP_HIP(hipEventRecord(t0, stream));
      if (off[J])
        hipLaunchKernelGGL(k_anchor_persist, dim3(grid), dim3(kBlock), 0, stream, d_jobs.p, d_off.p, J, d_head.p);
      MHGP_HIP(hipEventRecord(t1, stream));
      MHGP_LAUNCH_CHECK();
    } catch (const std::exception &e) {
      error = e.what();
    }
    launched = true;
    cv.notify_all();
  }
  void arrive(int slot, Context &c, const Params &P, uint32_t count) {
    hipEvent_t ev = nullptr;
    MHGP_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    MHGP_HIP(hipEventRecord(ev, c.stream));
    std::unique_lock<std::mutex> g(mu);
    if (slot < 0 || slot >= (int)done.size() || done[slot]) throw std::logic_error("batch slot reused");
    done[slot] = 1;
    jobs.push_back(P);
    counts.push_back(count);
    ready.push_back(ev);
    arrived++;
    if (arrived == expected) launch_locked();
    else cv.wait(g, [&] { return launched; });
    if (!error.empty()) throw HipError("batched anchor launch: " + error);
    g.unlock();
    MHGP_HIP(hipStreamWaitEvent(c.stream, t1, 0));
  }
  void leave(int slot) {
    std::lock_guard<std::mutex> g(mu);
    if (slot < 0 || slot >= (int)done.size() || done[slot]) return;
    done[slot] = 1;
    expected--;
    if (!launched && arrived > 0 && arrived == expected) launch_locked();
  }
};

std::shared_ptr<AnchorBatch> anchor_batch_create(int device, int participants) {
  auto b = std::make_shared<AnchorBatch>();
  b->device = device;
  b->expected = participants;
  b->done.assign(participants, 0);
  return b;
}

void anchor_batch_leave(AnchorBatch &b, int slot) { b.leave(slot); }

AnchorBatchStats anchor_batch_stats(AnchorBatch &b) {
  AnchorBatchStats st{};
  std::lock_guard<std::mutex> g(b.mu);
  st.jobs = (int64_t)b.jobs.size();
  st.anchors = b.anchors;
  if (b.launched && b.error.empty() && b.t1) {
    float ms = 0;
    MHGP_HIP(hipEventSynchronize(b.t1));
    MHGP_HIP(hipEventElapsedTime(&ms, b.t0, b.t1));
    st.ms = ms;
  }
  return st;
}

void supports_device(Context &c, int k_max, std::vector<EventRec> &events,
                     std::vector<FallbackRec> &fbs, std::vector<uint32_t> &fb_region,
                     mhgp_catalog_info *info) {
  const int64_t n = c.n;
  if (!c.catalog) c.catalog = std::make_unique<CatalogBuffers>();
  CatalogBuffers &B = *c.catalog;
  hipStream_t s = c.stream;
  const int64_t keff = std::min<int64_t>(k_max, n);
  const int smax = (int)std::min<int64_t>(keff + 1, n);
  B.k_max = k_max;
  B.s_max = smax;
  info->s_max = smax;
  if (n < 2 || smax < 2) {
    B.count = 0;  // no support of two or more points: an empty catalogue
    return;
  }
  if (smax > 11) throw std::invalid_argument("k_max must lie in 1..10");
  B.radius.ensure(n);
  B.gcount.ensure(n);
  B.goff.ensure(n);
  B.need.ensure(n);
  B.status.ensure(n);
  B.counters.ensure(8);
  B.anchors.ensure(n);
  if (B.cells.cap < 8 * kCells) {
    auto cells = make_cells();
    B.cells.ensure(8 * kCells);
    copy_wait(B.cells.p, cells.data(), cells.size() * 8, hipMemcpyHostToDevice, s);
    const std::vector<float> cf = make_cells_f(cells);
    B.cellf.ensure(8 * kCells);
    copy_wait(B.cellf.p, cf.data(), cf.size() * 4, hipMemcpyHostToDevice, s);
  }
  // Test hook (MHGP_TEST_ANCHOR_PATH): "region" sends every anchor through
  // the big path's probe and LDS region kernels, "slabs" through the
  // global-memory kernel, so that small clouds check those paths against the
  // exhaustive oracle.  Unset, the anchors take their normal tiers.
  int force_path = 0;
  if (const char *tp = std::getenv("MHGP_TEST_ANCHOR_PATH")) {
    if (std::strcmp(tp, "region") == 0) force_path = 1;
    else if (std::strcmp(tp, "slabs") == 0) force_path = 2;
  }
  // initial gather radius: twice the distance to the min(32, n-1)-th neighbour
  const int kq = (int)std::min<int64_t>(32, n - 1);
  StageTimer tm(s);
  DevBuf<uint32_t> &overflow = B.overflow;
  DevBuf<uint64_t> &total = B.total;
  overflow.ensure(1);
  total.ensure(1);
  MHGP_HIP(hipMemsetAsync(overflow.p, 0, 4, s));
  B.glist.ensure(1);
  const TileChoice tile = tiles_for(c, B, kq, overflow.p);
  launch_radius(tile.radius, s, c, kq, B);
  MHGP_LAUNCH_CHECK();
  std::vector<uint32_t> anchors(n);
  std::iota(anchors.begin(), anchors.end(), 0u);
  // anchors in Morton order: neighbouring workgroups share cache lines
  {
    std::vector<uint32_t> lid(n);
    copy_wait(lid.data(), c.leaf_id.p, n * 4, hipMemcpyDeviceToHost, s);
    anchors = lid;
  }
  double bounds[6];
  copy_wait(bounds, c.bounds.p, sizeof(bounds), hipMemcpyDeviceToHost, s);
  uint64_t event_cap = std::max<uint64_t>(1024, (uint64_t)n * (smax >= 11 ? 640 : 64 * smax));
  uint64_t fb_cap = std::max<uint64_t>(4096, (uint64_t)n / 4);
  uint64_t reg_cap = std::max<uint64_t>(1 << 20, (uint64_t)n * 64);
  int64_t widened = 0;
  int64_t open_anchors = 0, largest_open = 0;
  uint64_t ctr[8];
  // The record queues have a fixed capacity per launch; records past it are
  // dropped on the device and counted.  A catalogue that outgrows them runs
  // again with queues sized from the counts (the widened gather radii are
  // kept), unless the device cannot hold them.
  for (int attempt = 0;; ++attempt) {
  if (c.capacity_bytes > 0) {
    // the record queues are the catalogue's large, data-sized allocation
    const int64_t more = (int64_t)(event_cap - std::min<uint64_t>(event_cap, B.events.cap)) * (int64_t)sizeof(EventRec) +
                         (int64_t)(fb_cap - std::min<uint64_t>(fb_cap, B.fallbacks.cap)) * (int64_t)sizeof(FallbackRec) +
                         (int64_t)(reg_cap - std::min<uint64_t>(reg_cap, B.fb_region.cap)) * 4;
    if (device_bytes(c) + more > c.capacity_bytes)
      throw CapacityError("the catalogue's record queues (" + std::to_string(event_cap) + " records) exceed the " +
                          "context's device capacity of " + std::to_string(c.capacity_bytes) + " bytes");
  }
  B.events.ensure(event_cap);
  B.fallbacks.ensure(fb_cap);
  B.fb_region.ensure(reg_cap);
  MHGP_HIP(hipMemsetAsync(B.counters.p, 0, 8 * 8, s));
  MHGP_HIP(hipMemsetAsync(B.status.p, 0, n * 4, s));
  MHGP_HIP(hipMemsetAsync(overflow.p, 0, 4, s));
  open_anchors = 0;
  largest_open = 0;
  std::vector<uint32_t> pending = anchors;
  std::vector<uint32_t> big;
  for (int iter = 0; !pending.empty(); ++iter) {
    if (iter > 12) throw CapacityError("anchor gather radii did not converge");
    const uint32_t cnt = (uint32_t)pending.size();
    MHGP_HIP(hipMemcpyAsync(B.anchors.p, pending.data(), cnt * 4, hipMemcpyHostToDevice, s));
    // the first pass takes one thread per anchor (small radii); the widened
    // passes one wave per anchor
    if (iter == 0) launch_gather<false>(tile.gather, s, c, B, cnt, overflow.p);
    else launch_gather_wave<false>(s, c, B, cnt, overflow.p);
    hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, s, B.anchors.p, cnt, B.gcount.p,
                       B.goff.p, total.p);
    uint64_t tot = 0;
    MHGP_HIP(hipMemcpyAsync(&tot, total.p, 8, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipStreamSynchronize(s));
    B.glist.ensure(tot + 1);
    if (iter == 0) launch_gather<true>(tile.gather, s, c, B, cnt, overflow.p);
    else launch_gather_wave<true>(s, c, B, cnt, overflow.p);
    Params P{};
    P.X = c.X.p; P.Y = c.Y.p; P.Z = c.Z.p;
    P.gcount = B.gcount.p; P.goff = B.goff.p; P.glist = B.glist.p; P.radius = B.radius.p;
    P.cells = B.cells.p;
    P.cellf = B.cellf.p;
    for (int a = 0; a < 3; ++a) { P.blo[a] = bounds[a]; P.bhi[a] = bounds[3 + a]; }
    P.s_max = smax;
    P.need = B.need.p; P.status = B.status.p;
    P.events = B.events.p; P.event_cap = event_cap;
    P.fallbacks = B.fallbacks.p; P.fallback_cap = fb_cap;
    P.fb_region = B.fb_region.p; P.fb_region_cap = reg_cap;
    P.counters = (unsigned long long *)B.counters.p;
    P.anchors = B.anchors.p; P.anchor_count = cnt;
    P.big = 0; P.gcap = kGatherCap; P.rcap = kRegionCap;
    P.count_flops = c.count_flops ? 1 : 0;
    P.region_limit = std::min<uint32_t>(c.region_limit, kBigRegionMax);
    tm.mark("catalogue gather");
    uint64_t ctr_before[8];
    copy_wait(ctr_before, B.counters.p, sizeof(ctr_before), hipMemcpyDeviceToHost, s);
    if (c.batch && attempt == 0 && iter == 0) {
      // first pass of a batched cloud: one persistent launch over the batch
      c.batch->arrive(c.batch_slot, c, P, cnt);
    } else if (!force_path) {
    c.anchor_clock.start(s);
    hipLaunchKernelGGL(k_anchor_lds, dim3(cnt), dim3(kBlock), 0, s, P);
    c.anchor_clock.stop(s);
    MHGP_LAUNCH_CHECK();
    {
      uint64_t ctr_after[8];
      copy_wait(ctr_after, B.counters.p, sizeof(ctr_after), hipMemcpyDeviceToHost, s);
      const uint64_t ev = ctr_after[0] - ctr_before[0];
      c.stats.anchor_ms += c.anchor_clock.ms();
      c.stats.anchor_launches++;
      c.stats.gathered_points += (int64_t)tot;
      c.stats.events += (int64_t)ev;
      // algorithmic bytes: every gathered neighbour read once (id + FP64 xyz),
      // every record written once
      c.stats.anchor_bytes += (int64_t)tot * 28 + (int64_t)ev * (int64_t)sizeof(EventRec);
      // FP64 operations of the certified predicates the launch evaluated
      c.stats.anchor_flops += (int64_t)(ctr_after[5] - ctr_before[5]);
      c.stats.anchor_flops32 += (int64_t)(ctr_after[6] - ctr_before[6]);
    }
    }
    tm.mark("catalogue anchors");
    std::vector<uint32_t> st(n);
    std::vector<double> need(n), rad(n);
    std::vector<uint32_t> gc(n);
    MHGP_HIP(hipMemcpyAsync(st.data(), B.status.p, n * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(need.data(), B.need.p, n * 8, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(rad.data(), B.radius.p, n * 8, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipMemcpyAsync(gc.data(), B.gcount.p, n * 4, hipMemcpyDeviceToHost, s));
    MHGP_HIP(hipStreamSynchronize(s));
    if (force_path)
      for (uint32_t u : pending) st[u] = ST_BIG;
    std::vector<uint32_t> next, wide_list;
    for (int tier = 0; tier < 2; ++tier) {
      std::vector<uint32_t> wide;
      for (uint32_t u : (tier == 0 ? pending : wide_list)) {
        if (st[u] == ST_WIDEN) {
          rad[u] = std::max(need[u] * 1.02, rad[u] * 1.05);
          next.push_back(u);
          widened++;
        } else if (st[u] == ST_BIG) {
          if (tier == 0 && !force_path && gc[u] <= (uint32_t)kGatherCapWide) wide.push_back(u);
          else big.push_back(u);
        }
      }
      if (tier == 1 || wide.empty()) break;
      // anchors that outgrew the small LDS path: the wide one
      wide_list = wide;
      const uint32_t wc = (uint32_t)wide.size();
      MHGP_HIP(hipMemcpyAsync(B.anchors.p, wide.data(), wc * 4, hipMemcpyHostToDevice, s));
      Params W2 = P;
      W2.anchor_count = wc;
      W2.gcap = kGatherCapWide;
      W2.rcap = kRegionCapWide;
      hipLaunchKernelGGL(k_anchor_lds_wide, dim3(wc), dim3(kBlock), 0, s, W2);
      MHGP_LAUNCH_CHECK();
      MHGP_HIP(hipMemcpyAsync(st.data(), B.status.p, n * 4, hipMemcpyDeviceToHost, s));
      MHGP_HIP(hipMemcpyAsync(need.data(), B.need.p, n * 8, hipMemcpyDeviceToHost, s));
      MHGP_HIP(hipStreamSynchronize(s));
      if (tm.on) std::fprintf(stderr, "[mhgp] wide anchors: %u\n", wc);
      tm.mark("catalogue wide anchors");
    }
    // big anchors: widen inside the big path too
    if (!big.empty()) {
      std::vector<uint32_t> redo;
      for (int biter = 0; !big.empty(); ++biter) {
        if (biter > 12) throw std::runtime_error("big anchors did not converge");
        const uint32_t bc = (uint32_t)big.size();
        uint32_t gmax = 0;
        for (uint32_t u : big) gmax = std::max(gmax, gc[u]);
        gmax = std::max<uint32_t>(gmax, 64);
        // the CSR of the big anchors is still valid from this iteration's gather
        std::vector<uint32_t> again;
        // Launches in chunks of anchors of similar gather size (ascending), each
        // chunk's slabs sized by its own largest gather within one scratch
        // budget: a few anchors that gather most of the cloud no longer shrink
        // every launch to a handful of workgroups.  The budget (8 GiB, 1 GiB
        // in a context with a device capacity) keeps hundreds of workgroups
        // per launch even for gathers of 100 000 points (~10 MB of slab each).
        // The slabs are reused from one chunk to the next.
        uint64_t kScratchBudget = c.capacity_bytes > 0 ? (1ull << 30) : (8ull << 30);  // bytes
        if (c.capacity_bytes <= 0) {
          // and at most a quarter of the device memory still free (other
          // contexts on the device keep theirs), but never below 1 GiB
          size_t free_b = 0, total_b = 0;
          MHGP_HIP(hipMemGetInfo(&free_b, &total_b));
          kScratchBudget = std::max<uint64_t>(1ull << 30, std::min<uint64_t>(kScratchBudget, free_b / 4));
        }
        auto slab_of = [](uint32_t g) { return big_fc_offset(g) + 2ull * g; };
        auto launch_big = [&](const std::vector<uint32_t> &list_in, int probe) {
          std::vector<uint32_t> list(list_in);
          std::sort(list.begin(), list.end(),
                    [&](uint32_t x, uint32_t y) { return gc[x] < gc[y] || (gc[x] == gc[y] && x < y); });
          const size_t lc = list.size();
          std::vector<std::pair<size_t, uint32_t>> chunks;  // (end, gather cap)
          uint64_t scratch = 0;
          for (size_t b0 = 0; b0 < lc;) {
            size_t b1 = b0;
            uint32_t g = 64;
            while (b1 < lc) {
              const uint32_t gn = std::max<uint32_t>(64, gc[list[b1]]);
              if (b1 > b0 && slab_of(gn) * 8 * (b1 - b0 + 1) > kScratchBudget) break;
              g = gn;
              ++b1;
            }
            chunks.push_back({b1, g});
            scratch = std::max<uint64_t>(scratch, slab_of(g) * (b1 - b0));
            b0 = b1;
          }
          B.scratch.ensure(scratch);
          size_t b0 = 0;
          for (const auto &ch : chunks) {
            const uint32_t bn = (uint32_t)(ch.first - b0), g = ch.second;
            MHGP_HIP(hipMemcpyAsync(B.anchors.p, list.data() + b0, bn * 4, hipMemcpyHostToDevice, s));
            b0 = ch.first;
            Params Q = P;
            // the gather CSR may have been reallocated by a regather below
            Q.gcount = B.gcount.p; Q.goff = B.goff.p; Q.glist = B.glist.p; Q.radius = B.radius.p;
            Q.anchors = B.anchors.p; Q.anchor_count = bn; Q.big = 1; Q.probe = probe;
            Q.scratch = B.scratch.p; Q.slab_doubles = slab_of(g); Q.gcap = g; Q.rcap = g;
            hipLaunchKernelGGL(k_anchor_big, dim3(bn), dim3(kBlock), 0, s, Q);
            // probed regions that fit LDS: enumerated there, from this chunk's slabs
            if (probe && force_path != 2) {
              hipLaunchKernelGGL((k_anchor_region<kRegionCapWide, 0>), dim3(bn), dim3(kBlock), 0, s, Q);
              hipLaunchKernelGGL((k_anchor_region<kRegionCapHuge, kRegionCapWide>), dim3(bn), dim3(kBlock), 0, s, Q);
            }
            MHGP_LAUNCH_CHECK();
          }
          MHGP_HIP(hipMemcpyAsync(st.data(), B.status.p, n * 4, hipMemcpyDeviceToHost, s));
          MHGP_HIP(hipMemcpyAsync(need.data(), B.need.p, n * 8, hipMemcpyDeviceToHost, s));
          MHGP_HIP(hipStreamSynchronize(s));
        };
        // first the region sizes only (no enumeration), so that a region the
        // path cannot take is refused before any long launch
        launch_big(big, 1);
        tm.mark("catalogue big probe (+ LDS regions)");
        std::vector<uint32_t> full;
        for (uint32_t u : big) {
          if (st[u] == ST_WIDEN) {
            rad[u] = std::max(need[u] * 1.02, rad[u] * 1.05);
            again.push_back(u);
            widened++;
          } else if (st[u] == ST_HUGE) {
            // budget stop: the anchor stays open, the catalogue is conditional
            open_anchors++;
            largest_open = std::max<int64_t>(largest_open, (int64_t)need[u]);
          } else if (st[u] == ST_PROBED) {
            full.push_back(u);
          } else if (st[u] == ST_DONE) {
            // region enumerated in LDS after the probe
          } else {
            throw std::runtime_error("a big anchor did not complete its probe");
          }
        }
        if (tm.on)
          std::fprintf(stderr, "[mhgp] big pass %d: %u anchors, gather cap %u: %zu widen, %zu on slabs, %lld open so far\n",
                       biter, bc, gmax, again.size(), full.size(), (long long)open_anchors);
        if (!full.empty()) launch_big(full, 0);
        tm.mark("catalogue big on slabs");
        for (uint32_t u : full)
          if (st[u] != ST_DONE) throw std::runtime_error("a big anchor did not complete");
        if (again.empty()) break;
        // regather the widened big anchors
        const uint32_t ac = (uint32_t)again.size();
        copy_wait(B.radius.p, rad.data(), n * 8, hipMemcpyHostToDevice, s);
        MHGP_HIP(hipMemcpyAsync(B.anchors.p, again.data(), ac * 4, hipMemcpyHostToDevice, s));
        launch_gather_wave<false>(s, c, B, ac, overflow.p);
        hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, s, B.anchors.p, ac, B.gcount.p,
                           B.goff.p, total.p);
        MHGP_HIP(hipMemcpyAsync(&tot, total.p, 8, hipMemcpyDeviceToHost, s));
        MHGP_HIP(hipStreamSynchronize(s));
        B.glist.ensure(tot + 1);
        launch_gather_wave<true>(s, c, B, ac, overflow.p);
        MHGP_HIP(hipMemcpyAsync(gc.data(), B.gcount.p, n * 4, hipMemcpyDeviceToHost, s));
        MHGP_HIP(hipStreamSynchronize(s));
        tm.mark("catalogue big regather");
        big = again;
      }
      big.clear();
    }
    copy_wait(B.radius.p, rad.data(), n * 8, hipMemcpyHostToDevice, s);
    pending = next;
  }
  tm.mark("catalogue big anchors");
  uint32_t ovf = 0;
  copy_wait(&ovf, overflow.p, 4, hipMemcpyDeviceToHost, s);
  copy_wait(ctr, B.counters.p, sizeof(ctr), hipMemcpyDeviceToHost, s);
  if (ovf) throw std::runtime_error("LBVH traversal stack overflow in the gather");
  if (ctr[4]) throw std::runtime_error("face side queue overflow");
  if (ctr[0] <= event_cap && ctr[1] <= fb_cap && ctr[2] <= reg_cap) break;
  if (attempt >= 2) throw CapacityError("catalogue record queues did not settle after regrowth");
  const uint64_t ev2 = std::max(event_cap, ctr[0] + ctr[0] / 8), fb2 = std::max(fb_cap, ctr[1] + ctr[1] / 8),
                 rg2 = std::max(reg_cap, ctr[2] + ctr[2] / 8);
  size_t free_b = 0, total_b = 0;
  MHGP_HIP(hipMemGetInfo(&free_b, &total_b));
  const uint64_t more = (ev2 - std::min<uint64_t>(ev2, B.events.cap)) * sizeof(EventRec) +
                        (fb2 - std::min<uint64_t>(fb2, B.fallbacks.cap)) * sizeof(FallbackRec) +
                        (rg2 - std::min<uint64_t>(rg2, B.fb_region.cap)) * 4;
  if (more > free_b / 10 * 9)
    throw CapacityError("the catalogue needs " + std::to_string(ctr[0]) + " record slots, more than the device can hold");
  event_cap = ev2;
  fb_cap = fb2;
  reg_cap = rg2;
  // grown buffers start empty: release before the larger allocation
  if (B.events.cap < event_cap) B.events.release();
  if (B.fallbacks.cap < fb_cap) B.fallbacks.release();
  if (B.fb_region.cap < reg_cap) B.fb_region.release();
  }
  // records stay on the device; only the host's share of the work moves
  (void)events;
  B.count = (int64_t)ctr[0];
  fbs.resize(ctr[1]);
  fb_region.resize(ctr[2]);
  if (ctr[1]) copy_wait(fbs.data(), B.fallbacks.p, ctr[1] * sizeof(FallbackRec), hipMemcpyDeviceToHost, s);
  if (ctr[2]) copy_wait(fb_region.data(), B.fb_region.p, ctr[2] * 4, hipMemcpyDeviceToHost, s);
  tm.mark("catalogue download");
  info->candidate_count = (int64_t)ctr[3];
  info->anchors_widened = widened;
  info->anchors_open = open_anchors;
  info->largest_open_region = largest_open;
  info->status = open_anchors ? 1 : 0;
}

}  // namespace mhgp
