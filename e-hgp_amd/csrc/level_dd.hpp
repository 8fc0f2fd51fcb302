// Double-double squared radii with a running bound on their absolute error.
//
// The catalogue orders its records by FP64 levels whose error bounds are
// loose; neighbours whose FP64 intervals overlap are re-evaluated here with
// about 106 bits from the exact coordinate differences, and the order is
// certified when the double-double intervals are disjoint.  Every operation
// adds a bound on its own rounding error, taken generously above the
// published bounds of the double-word algorithms it uses (Joldes, Muller,
// Popescu, "Tight and rigorous error bounds for basic building blocks of
// double-word arithmetic", ACM TOMS 44(2), 2017: at most 3u^2 for the
// accurate sum and 7u^2 for the product, u = 2^-53), to the propagated error
// of its inputs.  Division bounds its rounding error a posteriori from the
// residual.  The error arithmetic itself is inflated by a relative 1e-12.
#pragma once

#include <cmath>

#include "exact_device.hpp"

namespace mhgp {

struct DDE {
  double hi, lo;  // value hi + lo
  double err;     // |exact - (hi + lo)| <= err (INFINITY: undefined)
};

constexpr double kU2 = 1.232595164407831e-32;  // 2^-106
constexpr double kErrUp = 1.0 + 1e-12;

__host__ __device__ inline void quick_two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
  s = a + b;
  e = b - (s - a);
}

__host__ __device__ inline double dd_mag(const DDE &a) { return fabs(a.hi) + fabs(a.lo); }

// exact difference a - b of two binary64 values
__host__ __device__ inline DDE dd_diff(double a, double b) {
  DDE r;
  two_diff(a, b, r.hi, r.lo);
  r.err = 0.0;
  return r;
}

__host__ __device__ inline DDE dd_neg(const DDE &a) { return {-a.hi, -a.lo, a.err}; }

__host__ __device__ inline DDE dd_scale(const DDE &a, double p2) {  // p2 a power of two
  return {a.hi * p2, a.lo * p2, a.err * fabs(p2)};
}

__host__ __device__ inline DDE dd_add(const DDE &a, const DDE &b) {
#pragma clang fp contract(off)
  double s, e, t, f;
  two_sum(a.hi, b.hi, s, e);
  two_sum(a.lo, b.lo, t, f);
  e += t;
  quick_two_sum(s, e, s, e);
  e += f;
  quick_two_sum(s, e, s, e);
  DDE r{s, e, 0.0};
  r.err = (a.err + b.err + 8.0 * kU2 * dd_mag(r)) * kErrUp;
  return r;
}

__host__ __device__ inline DDE dd_sub(const DDE &a, const DDE &b) { return dd_add(a, dd_neg(b)); }

__host__ __device__ inline DDE dd_mul(const DDE &a, const DDE &b) {
#pragma clang fp contract(off)
  double p, e;
  two_prod(a.hi, b.hi, p, e);
  const double t = a.hi * b.lo + a.lo * b.hi;
  e += t;
  quick_two_sum(p, e, p, e);
  DDE r{p, e, 0.0};
  r.err = (dd_mag(a) * b.err + dd_mag(b) * a.err + a.err * b.err + 16.0 * kU2 * dd_mag(r)) * kErrUp;
  return r;
}

__host__ __device__ inline DDE dd_sq(const DDE &a) { return dd_mul(a, a); }

// a / b; undefined when the interval of b contains zero
__host__ __device__ inline DDE dd_div(const DDE &a, const DDE &b) {
  const double bmag = fabs(b.hi + b.lo);
  if (!(bmag > 2.0 * b.err) || b.hi == 0.0) return {0.0, 0.0, INFINITY};
  const double q1 = a.hi / b.hi;
  const DDE bv{b.hi, b.lo, 0.0}, av{a.hi, a.lo, 0.0};
  DDE r = dd_sub(av, dd_mul(bv, {q1, 0.0, 0.0}));
  const double q2 = (r.hi + r.lo) / b.hi;
  double qh, ql;
  quick_two_sum(q1, q2, qh, ql);
  DDE q{qh, ql, 0.0};
  // rounding error of q from its residual a - q b (itself bounded)
  DDE res = dd_sub(av, dd_mul(bv, q));
  const double blow = bmag * (1.0 - 1e-15);
  const double round = (dd_mag(res) + res.err) / blow;
  // propagated error of the inputs
  const double prop = (a.err + dd_mag(q) * b.err) / (bmag - b.err);
  q.err = (round + prop) * kErrUp;
  return q;
}

__host__ __device__ inline DDE dd_dot3(const DDE *a, const DDE *b) {
  return dd_add(dd_add(dd_mul(a[0], b[0]), dd_mul(a[1], b[1])), dd_mul(a[2], b[2]));
}

__host__ __device__ inline void dd_cross(const DDE *a, const DDE *b, DDE *out) {
  out[0] = dd_sub(dd_mul(a[1], b[2]), dd_mul(a[2], b[1]));
  out[1] = dd_sub(dd_mul(a[2], b[0]), dd_mul(a[0], b[2]));
  out[2] = dd_sub(dd_mul(a[0], b[1]), dd_mul(a[1], b[0]));
}

// Squared circumradius of the support ids[0..m) (m = 2, 3 or 4), from the
// canonical coordinates: pair |a|^2/4; triangle |a|^2|b|^2|a-b|^2/(4D) with
// D = |a|^2|b|^2 - (a.b)^2; tetrahedron |N|^2/(4V^2) with
// N = |a|^2 (b x d) + |b|^2 (d x a) + |d|^2 (a x b), V = a.(b x d); a, b, d
// the exact differences to ids[0].
__host__ __device__ inline DDE dd_level(const double *X, const double *Y, const double *Z,
                                        const uint32_t *ids, int m) {
  if (m <= 1) return {0.0, 0.0, 0.0};  // a single point: level zero, exactly
  const uint32_t u = ids[0];
  DDE a[3], b[3], d[3];
  a[0] = dd_diff(X[ids[1]], X[u]);
  a[1] = dd_diff(Y[ids[1]], Y[u]);
  a[2] = dd_diff(Z[ids[1]], Z[u]);
  const DDE aa = dd_dot3(a, a);
  if (m == 2) return dd_scale(aa, 0.25);
  b[0] = dd_diff(X[ids[2]], X[u]);
  b[1] = dd_diff(Y[ids[2]], Y[u]);
  b[2] = dd_diff(Z[ids[2]], Z[u]);
  const DDE bb = dd_dot3(b, b);
  if (m == 3) {
    DDE c[3];
    c[0] = dd_diff(X[ids[2]], X[ids[1]]);
    c[1] = dd_diff(Y[ids[2]], Y[ids[1]]);
    c[2] = dd_diff(Z[ids[2]], Z[ids[1]]);
    const DDE ab = dd_dot3(a, b);
    const DDE D = dd_sub(dd_mul(aa, bb), dd_sq(ab));
    const DDE num = dd_mul(dd_mul(aa, bb), dd_dot3(c, c));
    return dd_scale(dd_div(num, D), 0.25);
  }
  d[0] = dd_diff(X[ids[3]], X[u]);
  d[1] = dd_diff(Y[ids[3]], Y[u]);
  d[2] = dd_diff(Z[ids[3]], Z[u]);
  const DDE dd = dd_dot3(d, d);
  DDE bxd[3], dxa[3], axb[3], N[3];
  dd_cross(b, d, bxd);
  dd_cross(d, a, dxa);
  dd_cross(a, b, axb);
  const DDE V = dd_dot3(a, bxd);
  for (int k = 0; k < 3; ++k) N[k] = dd_add(dd_add(dd_mul(aa, bxd[k]), dd_mul(bb, dxa[k])), dd_mul(dd, axb[k]));
  return dd_scale(dd_div(dd_dot3(N, N), dd_sq(V)), 0.25);
}

// +1 when x < y is certain, -1 when y < x is certain, 0 otherwise
__host__ __device__ inline int dd_order(const DDE &x, const DDE &y) {
  if (!(x.err < INFINITY) || !(y.err < INFINITY)) return 0;
  const DDE diff = dd_sub({y.hi, y.lo, 0.0}, {x.hi, x.lo, 0.0});
  const double bound = (x.err + y.err + diff.err) * kErrUp;
  const double v = diff.hi + diff.lo;
  if (v > bound) return 1;
  if (-v > bound) return -1;
  return 0;
}

__host__ __device__ inline bool dd_less(const DDE &x, const DDE &y) {
  return x.hi < y.hi || (x.hi == y.hi && x.lo < y.lo);
}

}  // namespace mhgp
