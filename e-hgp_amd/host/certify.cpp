// Exact re-certification of an LBVH snapshot (host, no device).
//
// Role of the reference's MortonLbvhIndex::import_certified_snapshot
// (morsehgp3d/src/cpu/spatial/lbvh.cpp:562-829): a snapshot proposed by any
// device is accepted only when every field is the exact Morton LBVH of the
// canonical cloud.  Checked here:
//  * the cloud is canonical: finite binary64 coordinates, strictly increasing
//    in the lexicographic total order of their keys (-0 equals +0);
//  * the global frame: per axis the least and greatest coordinate;
//  * every leaf code is the interleave of q = floor(2^21 (x - l) / (u - l))
//    (2^21 - 1 at x = u, 0 on a flat axis), verified exactly as
//    q (u - l) <= 2^21 (x - l) < (q + 1) (u - l) on the integers of the
//    dyadic cloud;
//  * the leaves are the PointIds ordered by (code, PointId);
//  * the topology: 2n - 1 nodes in strict postorder, each internal node over
//    leaves [b, e) split at the first leaf whose code has the highest bit in
//    which the codes of leaves b and e - 1 differ, or at b + (e - b) / 2 when
//    they are equal; leaves have no children and a one-leaf range;
//  * witnesses: per axis the PointId of the least / greatest coordinate of the
//    node's points, ties to the smaller PointId;
//  * counters: point and node count, maximum depth, the number of runs of two
//    or more equal codes and the longest such run.
#include "certify.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "exact.hpp"

namespace mhgp {

namespace {

uint64_t key_of(double x) {
  uint64_t b;
  std::memcpy(&b, &x, 8);
  if (b == 0x8000000000000000ull) b = 0;
  return (b >> 63) ? ~b : (b ^ 0x8000000000000000ull);
}

uint64_t spread3(uint64_t q[3]) {
  uint64_t c = 0;
  for (int bit = 0; bit < 21; ++bit)
    for (int a = 0; a < 3; ++a) c |= ((q[a] >> bit) & 1ull) << (3 * bit + 2 - a);
  return c;
}

void unspread3(uint64_t code, uint64_t q[3]) {
  q[0] = q[1] = q[2] = 0;
  for (int bit = 0; bit < 21; ++bit)
    for (int a = 0; a < 3; ++a) q[a] |= ((code >> (3 * bit + 2 - a)) & 1ull) << bit;
}

struct Fail {
  std::string what;
};

}  // namespace

std::string certify_snapshot(const double *xyz, int64_t n, const uint64_t *leaves, const uint64_t *nodes,
                             const uint64_t *counters) {
  try {
    if (n <= 0) throw Fail{"empty cloud"};
    const int64_t m = 2 * n - 1;
    std::vector<uint64_t> key(3 * (size_t)n);
    for (int64_t i = 0; i < 3 * n; ++i) {
      uint64_t b;
      std::memcpy(&b, &xyz[i], 8);
      if (((b >> 52) & 0x7ff) == 0x7ff) throw Fail{"non-finite coordinate"};
      key[i] = key_of(xyz[i]);
    }
    for (int64_t i = 1; i < n; ++i) {
      const uint64_t *p = &key[3 * (i - 1)], *q = &key[3 * i];
      const bool less = p[0] < q[0] || (p[0] == q[0] && (p[1] < q[1] || (p[1] == q[1] && p[2] < q[2])));
      if (!less) throw Fail{"the cloud is not canonical at PointId " + std::to_string(i)};
    }
    // frame: least / greatest coordinate per axis (ties to the smaller PointId)
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int64_t i = 1; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        if (key[3 * i + a] < key[3 * lo[a] + a]) lo[a] = i;
        if (key[3 * i + a] > key[3 * hi[a] + a]) hi[a] = i;
      }
    ExactCloud ex;
    ex.load(xyz, n);
    BigInt ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = ex.coord(hi[a], a) - ex.coord(lo[a], a);
    const BigInt two21 = BigInt::from_dyadic(1, 21);
    // leaves: a permutation in (code, PointId) order with exact codes
    std::vector<char> seen(n, 0);
    for (int64_t i = 0; i < n; ++i) {
      const uint64_t code = leaves[2 * i], id = leaves[2 * i + 1];
      if (id >= (uint64_t)n || seen[id]) throw Fail{"leaf " + std::to_string(i) + " repeats or leaves the cloud"};
      seen[id] = 1;
      if (i && (leaves[2 * (i - 1)] > code || (leaves[2 * (i - 1)] == code && leaves[2 * i - 1] >= id)))
        throw Fail{"leaves are not in (code, PointId) order at " + std::to_string(i)};
      if (code >> 63) throw Fail{"a code uses bit 63"};
      uint64_t q[3];
      unspread3(code, q);
      if (spread3(q) != code) throw Fail{"a code is not an interleave"};
      for (int a = 0; a < 3; ++a) {
        if (ext[a].sign() == 0) {
          if (q[a] != 0) throw Fail{"a flat axis has a nonzero bin"};
          continue;
        }
        const BigInt &x = ex.coord((int64_t)id, a), &l = ex.coord(lo[a], a);
        if (compare(x, ex.coord(hi[a], a)) == 0) {
          if (q[a] != (1ull << 21) - 1) throw Fail{"the upper bound's bin is not 2^21 - 1"};
          continue;
        }
        const BigInt scaled = (x - l) * two21, qa((int64_t)q[a]), qb((int64_t)q[a] + 1);
        if (compare(qa * ext[a], scaled) > 0 || compare(scaled, qb * ext[a]) >= 0)
          throw Fail{"the Morton code of PointId " + std::to_string(id) + " is not its exact quantisation"};
      }
    }
    // topology, rebuilt from the certified leaves in strict postorder
    struct Frame {
      uint64_t b, e, depth;
      int state;
      uint64_t left;
    };
    std::vector<Frame> stack{{0, (uint64_t)n, 0, 0, 0}};
    std::vector<uint64_t> ret;
    int64_t next = 0;
    uint64_t maxdepth = 0;
    auto lower = [&](uint64_t x, uint64_t y, int a) {
      const uint64_t kx = key[3 * x + a], ky = key[3 * y + a];
      return (ky < kx || (ky == kx && y < x)) ? y : x;
    };
    auto upper = [&](uint64_t x, uint64_t y, int a) {
      const uint64_t kx = key[3 * x + a], ky = key[3 * y + a];
      return (ky > kx || (ky == kx && y < x)) ? y : x;
    };
    auto split = [&](uint64_t b, uint64_t e) {
      const uint64_t fc = leaves[2 * b], lc = leaves[2 * (e - 1)];
      if (fc == lc) return b + (e - b) / 2;
      const uint64_t bit = 1ull << (63 - __builtin_clzll(fc ^ lc));
      uint64_t s = b + 1;
      while (s < e && !(leaves[2 * s] & bit)) ++s;
      return s;
    };
    while (!stack.empty()) {
      Frame &f = stack.back();
      if (f.depth > maxdepth) maxdepth = f.depth;
      if (next >= m) throw Fail{"more nodes than 2n - 1"};
      if (f.e - f.b == 1) {
        const uint64_t *nd = nodes + 10 * next;
        const uint64_t pid = leaves[2 * f.b + 1];
        for (int a = 0; a < 6; ++a)
          if (nd[a] != pid) throw Fail{"leaf node " + std::to_string(next) + " has a wrong witness"};
        if (nd[6] != ~0ull || nd[7] != ~0ull || nd[8] != f.b || nd[9] != f.e)
          throw Fail{"leaf node " + std::to_string(next) + " is malformed"};
        ret.push_back((uint64_t)next++);
        stack.pop_back();
        continue;
      }
      const uint64_t s = split(f.b, f.e);
      if (f.state == 0) {
        f.state = 1;
        const uint64_t b = f.b, d = f.depth + 1;
        stack.push_back({b, s, d, 0, 0});
        continue;
      }
      if (f.state == 1) {
        f.state = 2;
        f.left = ret.back();
        ret.pop_back();
        const uint64_t e = f.e, d = f.depth + 1;
        stack.push_back({s, e, d, 0, 0});
        continue;
      }
      const uint64_t right = ret.back();
      ret.pop_back();
      const uint64_t *nd = nodes + 10 * next, *L = nodes + 10 * f.left, *R = nodes + 10 * right;
      if (nd[6] != f.left || nd[7] != right || nd[8] != f.b || nd[9] != f.e)
        throw Fail{"node " + std::to_string(next) + " does not follow the split rule or the postorder"};
      for (int a = 0; a < 3; ++a)
        if (nd[a] != lower(L[a], R[a], a) || nd[3 + a] != upper(L[3 + a], R[3 + a], a))
          throw Fail{"node " + std::to_string(next) + " has a wrong witness"};
      ret.push_back((uint64_t)next++);
      stack.pop_back();
    }
    if (next != m) throw Fail{"fewer nodes than 2n - 1"};
    uint64_t groups = 0, longest = 0;
    for (int64_t b = 0; b < n;) {
      int64_t e = b + 1;
      while (e < n && leaves[2 * e] == leaves[2 * b]) ++e;
      if (e - b > 1) {
        groups++;
        longest = std::max<uint64_t>(longest, (uint64_t)(e - b));
      }
      b = e;
    }
    if (counters[0] != (uint64_t)n || counters[1] != (uint64_t)m || counters[2] != maxdepth ||
        counters[3] != groups || counters[4] != longest)
      throw Fail{"the snapshot counters differ from the certified ones"};
    return std::string();
  } catch (const Fail &f) {
    return f.what;
  }
}

}  // namespace mhgp
