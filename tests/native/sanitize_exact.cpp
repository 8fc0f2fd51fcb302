// Sanitizer driver of the host-only exact layers (host/exact.cpp,
// host/certify.cpp), built with g++ under AddressSanitizer + UBSan and under
// ThreadSanitizer by tests/test_host_sanitizers.py.  Every check compares the
// multiprecision code with 128-bit integer arithmetic on small integer
// clouds, where the same quantities fit, or with an identity of the
// mathematics (the lifted insphere determinant, the circumcentre of a
// triangle in its plane), so that a sanitizer run also exercises real work:
//   arith        BigInt +, -, *, compare, dyadic shifts, conversions
//   supports T   ExactSupport analyze / side / level on T threads sharing one
//                lazily converted ExactCloud (the host exact run's pattern)
//   certify F    certify_snapshot on the snapshot in file F (int64 n, then
//                xyz 3n f64, leaves 2n u64, nodes 10(2n-1) u64, counters 5
//                u64): accepted as given, rejected with one leaf code changed
// Prints "ok" and exits 0, or names the first failed check and exits 1.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../e-hgp_amd/host/certify.hpp"
#include "../../e-hgp_amd/host/exact.hpp"

using namespace mhgp;
using i128 = __int128;

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  int64_t range(int64_t lo, int64_t hi) { return (int64_t)((uint64_t)lo + next() % ((uint64_t)hi - (uint64_t)lo + 1)); }
};

int fails = 0;
#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      if (fails++ < 10) {                                  \
        std::fprintf(stderr, "check failed: %s: ", #cond); \
        std::fprintf(stderr, __VA_ARGS__);                 \
        std::fprintf(stderr, "\n");                        \
      }                                                    \
    }                                                      \
  } while (0)

int sgn(i128 v) { return (v > 0) - (v < 0); }

// v = h 2^62 + l with 0 <= l < 2^62 (arithmetic shift), for any 128-bit v
BigInt big(i128 v) {
  int64_t h = (int64_t)(v >> 62);
  int64_t l = (int64_t)(v & (((i128)1 << 62) - 1));
  return BigInt(h) * BigInt::from_dyadic(1, 62) + BigInt(l);
}

void arith() {
  Rng r{1};
  const int64_t L = int64_t(1) << 62;
  for (int it = 0; it < 20000; ++it) {
    int64_t a = r.range(-L, L), b = r.range(-L, L), c = r.range(-L, L), d = r.range(-L, L);
    BigInt A(a), B(b), C(c), D(d);
    CHECK(compare(A * B, C * D) == sgn((i128)a * b - (i128)c * d), "products %d", it);
    CHECK(compare(A + B, C) == sgn((i128)a + b - c), "sum %d", it);
    CHECK(compare(A - B, -(B - A)) == 0, "negation %d", it);
    CHECK(compare(big((i128)a * b), A * B) == 0, "split constructor %d", it);
    i128 p = (i128)a * b;
    double want = (double)p, got = (A * B).to_double();
    CHECK(std::fabs(got - want) <= std::fabs(want) * 0x1p-52, "to_double %d", it);
    int sh = 0;
    double top = (A * B).top_bits(sh);
    CHECK(std::fabs(std::ldexp(top, sh) - want) <= std::fabs(want) * 0x1p-52, "top_bits %d", it);
    // multi-limb: the same six factors in two orders, then an add/subtract round trip
    BigInt X = ((A * B) * (C * D)) * (A * C);
    BigInt Y = (A * (A * C)) * ((D * B) * C);
    CHECK(compare(X, Y) == 0, "commuted product %d", it);
    CHECK(compare((X + Y) - Y, X) == 0, "add/sub round trip %d", it);
    CHECK(X.sign() == (a != 0 && c != 0 ? sgn(b) * sgn(d) : 0), "product sign %d", it);
  }
  for (int k = 0; k < 200; ++k) {
    int64_t s = r.range(-1000000, 1000000);
    CHECK(compare(BigInt::from_dyadic(s, k) * BigInt(2), BigInt::from_dyadic(s, k + 1)) == 0, "dyadic %d", k);
  }
  CHECK(compare(BigInt::from_dyadic(12345, 0), BigInt(12345)) == 0, "dyadic 0");
  for (int it = 0; it < 5000; ++it) {
    int64_t n1 = r.range(1, L), d1 = r.range(1, L), n2 = r.range(1, L), d2 = r.range(1, L);
    if (it % 7 == 0) { n2 = n1; d2 = d1; }
    CHECK(compare_levels(BigInt(n1), BigInt(d1), BigInt(n2), BigInt(d2)) == sgn((i128)n1 * d2 - (i128)n2 * d1),
          "compare_levels %d", it);
    int e = (int)r.range(-60, 60);
    double want = std::ldexp((double)((long double)n1 / (long double)d1), e);
    double got = ratio_to_double(BigInt(n1), BigInt(d1), e);
    CHECK(std::fabs(got - want) <= std::fabs(want) * 1e-15, "ratio_to_double %d", it);
  }
}

struct P3 { int64_t x, y, z; };
P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
i128 dot(P3 a, P3 b) { return (i128)a.x * b.x + (i128)a.y * b.y + (i128)a.z * b.z; }
P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

i128 det3(P3 a, P3 b, P3 c) { return dot(a, cross(b, c)); }

// det of rows (p - x, |p - x|^2) for the four points: equals f(x) * orient(a, b, c, d)
// with f(x) = |x - centre|^2 - r^2 and orient = det[b - a; c - a; d - a]
i128 insphere(const P3 *q, P3 x) {
  i128 m[4][4];
  for (int i = 0; i < 4; ++i) {
    P3 d = sub(q[i], x);
    m[i][0] = d.x; m[i][1] = d.y; m[i][2] = d.z; m[i][3] = dot(d, d);
  }
  auto minor3 = [&](int skip_col, int r0, int r1, int r2) {
    int c[3], k = 0;
    for (int j = 0; j < 4; ++j) if (j != skip_col) c[k++] = j;
    auto e = [&](int r, int j) { return m[r][c[j]]; };
    return e(r0, 0) * (e(r1, 1) * e(r2, 2) - e(r1, 2) * e(r2, 1)) -
           e(r0, 1) * (e(r1, 0) * e(r2, 2) - e(r1, 2) * e(r2, 0)) +
           e(r0, 2) * (e(r1, 0) * e(r2, 1) - e(r1, 1) * e(r2, 0));
  };
  // expansion along row 0
  i128 det = 0;
  for (int j = 0; j < 4; ++j) {
    i128 t = m[0][j] * minor3(j, 1, 2, 3);
    det += (j % 2 == 0) ? t : -t;
  }
  return det;
}

void supports_worker(const ExactCloud &cloud, const std::vector<P3> &pts, uint64_t seed, int rounds) {
  Rng r{seed};
  int64_t n = (int64_t)pts.size();
  for (int it = 0; it < rounds; ++it) {
    int m = 2 + (int)(r.next() % 3);
    int64_t id[4];
    for (int k = 0; k < m; ++k) {
      bool again;
      do {
        id[k] = (int64_t)(r.next() % (uint64_t)n);
        again = false;
        for (int j = 0; j < k; ++j) again |= id[j] == id[k];
      } while (again);
    }
    ExactSupport s;
    s.analyze(cloud, id, m);
    P3 q[4];
    for (int k = 0; k < m; ++k) q[k] = pts[id[k]];
    if (m == 2) {
      i128 dd = dot(sub(q[0], q[1]), sub(q[0], q[1]));
      CHECK(!s.dependent && s.well, "pair flags");
      CHECK(compare_levels(s.lnum, s.lden, big(dd), BigInt(4)) == 0, "pair level");
    } else if (m == 3) {
      P3 u = sub(q[1], q[0]), v = sub(q[2], q[0]), nn = cross(u, v);
      i128 n2 = dot(nn, nn);
      CHECK(s.dependent == (n2 == 0), "triangle dependence");
      if (n2 == 0) continue;
      i128 du = dot(u, u), dv = dot(v, v);
      P3 a1 = cross(v, nn), a2 = cross(nn, u);
      // centre - q0 = N / (2 |n|^2), N = |u|^2 (v x n) + |v|^2 (n x u)
      i128 N[3] = {du * a1.x + dv * a2.x, du * a1.y + dv * a2.y, du * a1.z + dv * a2.z};
      i128 NN = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
      CHECK(compare_levels(s.lnum, s.lden, big(NN), big(4 * n2 * n2)) == 0, "triangle level");
      i128 ab = dot(u, v), bc = dot(sub(q[0], q[1]), sub(q[2], q[1])), ca = dot(sub(q[0], q[2]), sub(q[1], q[2]));
      if (ab > 0 && bc > 0 && ca > 0) CHECK(s.well, "acute triangle is well centred");
      if (ab < 0 || bc < 0 || ca < 0) CHECK(!s.well, "obtuse triangle is not well centred");
      for (int t = 0; t < 8; ++t) {
        int64_t x = (int64_t)(r.next() % (uint64_t)n);
        P3 w = sub(pts[x], q[0]);
        i128 W[3] = {2 * n2 * w.x, 2 * n2 * w.y, 2 * n2 * w.z};
        i128 e0 = W[0] - N[0], e1 = W[1] - N[1], e2 = W[2] - N[2];
        int want = sgn(e0 * e0 + e1 * e1 + e2 * e2 - NN);
        CHECK(s.side(cloud, x) == want, "triangle side");
      }
      for (int k = 0; k < 3; ++k) CHECK(s.side(cloud, id[k]) == 0, "triangle vertex on its sphere");
    } else {
      i128 o = det3(sub(q[1], q[0]), sub(q[2], q[0]), sub(q[3], q[0]));
      CHECK(s.dependent == (o == 0), "tetrahedron dependence");
      if (o == 0) continue;
      for (int t = 0; t < 8; ++t) {
        int64_t x = (int64_t)(r.next() % (uint64_t)n);
        CHECK(s.side(cloud, x) == sgn(insphere(q, pts[x])) * sgn(o), "tetrahedron side");
      }
      for (int k = 0; k < 4; ++k) CHECK(s.side(cloud, id[k]) == 0, "tetrahedron vertex on its sphere");
    }
    // distances agree with the integers whatever thread converted the points
    int64_t a = id[0], b = id[1], c = (int64_t)(r.next() % (uint64_t)n);
    i128 da = dot(sub(pts[c], pts[a]), sub(pts[c], pts[a])), db = dot(sub(pts[c], pts[b]), sub(pts[c], pts[b]));
    CHECK(cloud.cmp_dist(c, a, b) == sgn(da - db), "cmp_dist");
  }
}

void supports(int threads) {
  // small integer grid: exact ties (cospherical and coplanar sets) are common;
  // the point (1, 1, 1) fixes the cloud's common exponent at 0
  Rng r{7};
  std::vector<P3> pts;
  pts.push_back({1, 1, 1});
  for (int i = 0; i < 1500; ++i) pts.push_back({r.range(-20, 20), r.range(-20, 20), r.range(-20, 20)});
  std::vector<double> xyz;
  for (auto &p : pts) { xyz.push_back((double)p.x); xyz.push_back((double)p.y); xyz.push_back((double)p.z); }
  ExactCloud cloud;
  cloud.load(xyz.data(), (int64_t)pts.size());
  CHECK(cloud.e0() == 0, "common exponent");
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back(supports_worker, std::cref(cloud), std::cref(pts), 100 + (uint64_t)t, threads > 1 ? 400 : 3000);
  for (auto &t : pool) t.join();
}

int certify(const char *path) {
  FILE *f = std::fopen(path, "rb");
  if (!f) { std::fprintf(stderr, "cannot open %s\n", path); return 1; }
  int64_t n = 0;
  bool ok = std::fread(&n, 8, 1, f) == 1 && n > 0 && n < (int64_t(1) << 24);
  std::vector<double> xyz(ok ? 3 * n : 0);
  std::vector<uint64_t> leaves(ok ? 2 * n : 0), nodes(ok ? 10 * (2 * n - 1) : 0), counters(5);
  ok = ok && std::fread(xyz.data(), 8, xyz.size(), f) == xyz.size() &&
       std::fread(leaves.data(), 8, leaves.size(), f) == leaves.size() &&
       std::fread(nodes.data(), 8, nodes.size(), f) == nodes.size() && std::fread(counters.data(), 8, 5, f) == 5;
  std::fclose(f);
  if (!ok) { std::fprintf(stderr, "short snapshot file\n"); return 1; }
  std::string why = certify_snapshot(xyz.data(), n, leaves.data(), nodes.data(), counters.data());
  CHECK(why.empty(), "snapshot rejected: %s", why.c_str());
  leaves[2 * (n / 2)] ^= 1;
  CHECK(!certify_snapshot(xyz.data(), n, leaves.data(), nodes.data(), counters.data()).empty(),
        "changed leaf code accepted");
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "arith") arith();
  else if (mode == "supports" && argc > 2) supports(std::atoi(argv[2]));
  else if (mode == "certify" && argc > 2) { if (certify(argv[2])) return 1; }
  else { std::fprintf(stderr, "usage: %s arith | supports THREADS | certify FILE\n", argv[0]); return 2; }
  if (fails) { std::fprintf(stderr, "%d checks failed\n", fails); return 1; }
  std::printf("ok\n");
  return 0;
}
