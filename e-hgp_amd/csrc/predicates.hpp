// Certified FP64 sphere predicates for supports of 2, 3 and 4 points.
//
// Every predicate is a division-free integer-coefficient polynomial in the
// coordinates relative to the support's first point (the anchor).  It is
// evaluated twice in FP64: once signed (P) and once with every subtraction
// replaced by an addition of magnitudes (Pabs).  With every input and every
// operation rounded to nearest, |fl(P) - P| <= k u Pabs for k operations on
// the longest chain; k u stays below 1e-14 for these polynomials, so a sign
// is certified when |fl(P)| > kCert * Pabs + kCertAbs (kCert = 1e-12; the
// absolute term covers underflow) and left to the exact host path otherwise.
//
// Formulas (a, b, d offsets of the other support points, p of the query):
//  pair:        P = p.(p - a)                         (diametral sphere)
//  triangle:    D = aa bb - ab^2,  al = bb (aa - ab),  be = aa (bb - ab)
//               well-centred <=> al > 0, be > 0, 2D - al - be > 0
//               P = D pp - al pa - be pb = D pp - p.g,  g = al a + be b
//               (the bound of p.g is sum |p_i| (|al a_i| + |be b_i|), the
//               same magnitude polynomial as for the expanded form)
//               r^2 = aa bb |a - b|^2 / (4 D)
//  tetrahedron: V = a.(b x d),  N = aa (b x d) + bb (d x a) + dd (a x b)
//               well-centred <=> N.(b x d), N.(d x a), N.(a x b) > 0 and
//               2 V^2 - N.(b x d) - N.(d x a) - N.(a x b) > 0
//               P = V (V pp - p.N),  r^2 = |N|^2 / (4 V^2)
// (circumcentre in the affine hull and barycentric well-centring as in the
// reference's exact/center.hpp:263-340 and exact/support.hpp:41-135).
#pragma once

#include <hip/hip_runtime.h>

namespace mhgp {

constexpr double kCert = 1e-12;

// Absolute part of the error bound.  Once partial products reach the
// subnormal range each rounding adds up to an absolute 2^-1074 (and up to
// 2^-1022 where a unit flushes subnormals), which no relative bound covers;
// the polynomials here take fewer than 200 roundings, so kCertAbs exceeds
// their total absolute error with a wide margin.  Values below it are left
// undecided, i.e. to the exact path.
constexpr double kCertAbs = 1e-300;

// -1 / +1 when certain, 0 when undecided
__host__ __device__ inline int cert_sign(double v, double vabs) {
  const double e = kCert * vabs + kCertAbs;
  return v > e ? 1 : (v < -e ? -1 : 0);
}

struct V3 {
  double x, y, z;
};
__host__ __device__ inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline double adot(const V3 &a, const V3 &b) {
  return fabs(a.x * b.x) + fabs(a.y * b.y) + fabs(a.z * b.z);
}
__host__ __device__ inline V3 cross(const V3 &a, const V3 &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline V3 across(const V3 &a, const V3 &b) {
  return {fabs(a.y * b.z) + fabs(a.z * b.y), fabs(a.z * b.x) + fabs(a.x * b.z),
          fabs(a.x * b.y) + fabs(a.y * b.x)};
}
__host__ __device__ inline V3 vabs(const V3 &a) { return {fabs(a.x), fabs(a.y), fabs(a.z)}; }
__host__ __device__ inline V3 sub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

struct PairSphere {
  V3 a;
  __host__ __device__ void init(const V3 &aa) { a = aa; }
  __host__ __device__ int side(const V3 &p) const {
    V3 q = sub(p, a);
    double v = dot(p, q);
    double va = fabs(p.x) * (fabs(p.x) + fabs(a.x)) + fabs(p.y) * (fabs(p.y) + fabs(a.y)) +
                fabs(p.z) * (fabs(p.z) + fabs(a.z));
    return cert_sign(v, va);
  }
  __host__ __device__ double level() const { return 0.25 * dot(a, a); }
};

struct TriSphere {
  V3 a, b, n, nA, g, gA;
  double aa, bb, ab, abA, D, DA, al, alA, be, beA;
  // returns +1 certainly well-centred, -1 certainly not or dependent, 0 undecided
  __host__ __device__ int init(const V3 &va, const V3 &vb) {
    a = va;
    b = vb;
    aa = dot(a, a);
    bb = dot(b, b);
    ab = dot(a, b);
    abA = adot(a, b);
    D = aa * bb - ab * ab;
    DA = aa * bb + abA * abA;
    al = bb * (aa - ab);
    alA = bb * (aa + abA);
    be = aa * (bb - ab);
    beA = aa * (bb + abA);
    n = cross(a, b);
    nA = across(a, b);
    g = {al * a.x + be * b.x, al * a.y + be * b.y, al * a.z + be * b.z};
    gA = {alA * fabs(a.x) + beA * fabs(b.x), alA * fabs(a.y) + beA * fabs(b.y), alA * fabs(a.z) + beA * fabs(b.z)};
    int sd = cert_sign(D, DA);
    if (sd <= 0) return sd < 0 ? -1 : 0;  // D >= 0 always; 0 is dependent or undecided
    int s1 = cert_sign(al, alA), s2 = cert_sign(be, beA);
    int s3 = cert_sign(2 * D - al - be, 2 * DA + alA + beA);
    if (s1 < 0 || s2 < 0 || s3 < 0) return -1;
    if (s1 == 0 || s2 == 0 || s3 == 0) return 0;
    return 1;
  }
  __host__ __device__ bool nondegenerate() const { return cert_sign(D, DA) > 0; }
  // scaled power P = D pp - p.g and its magnitude bound
  __host__ __device__ void power(const V3 &p, double &P, double &PA) const {
    const double pp = dot(p, p);
    P = D * pp - dot(p, g);
    PA = DA * pp + adot(p, gA);
  }
  __host__ __device__ int side(const V3 &p) const {
    double P, PA;
    power(p, P, PA);
    return cert_sign(P, PA);
  }
  // signed height H = p.(a x b) and its bound
  __host__ __device__ void height(const V3 &p, double &H, double &HA) const {
    H = dot(p, n);
    HA = adot(p, nA);
  }
  __host__ __device__ double level() const {
    V3 c = sub(a, b);
    return aa * bb * dot(c, c) / (4.0 * D);
  }
  __host__ __device__ double level_err() const { return 8.0 * kCert * (1.0 + DA / D); }
  __host__ __device__ double r2() const { return level(); }
};

struct TetSphere {
  V3 a, b, d, N, NA;
  double V, VA;
  __host__ __device__ int init(const V3 &va, const V3 &vb, const V3 &vd) {
    a = va;
    b = vb;
    d = vd;
    V3 bxd = cross(b, d), dxa = cross(d, a), axb = cross(a, b);
    V3 bxdA = across(b, d), dxaA = across(d, a), axbA = across(a, b);
    V = dot(a, bxd);
    VA = adot(vabs(a), bxdA);
    if (cert_sign(V, VA) == 0) return 0;
    double aa = dot(a, a), bb = dot(b, b), dd = dot(d, d);
    N = {aa * bxd.x + bb * dxa.x + dd * axb.x, aa * bxd.y + bb * dxa.y + dd * axb.y,
         aa * bxd.z + bb * dxa.z + dd * axb.z};
    NA = {aa * bxdA.x + bb * dxaA.x + dd * axbA.x, aa * bxdA.y + bb * dxaA.y + dd * axbA.y,
          aa * bxdA.z + bb * dxaA.z + dd * axbA.z};
    double nv = dot(N, bxd), nw = dot(N, dxa), nz = dot(N, axb);
    double nvA = dot(NA, bxdA), nwA = dot(NA, dxaA), nzA = dot(NA, axbA);
    int s1 = cert_sign(nv, nvA), s2 = cert_sign(nw, nwA), s3 = cert_sign(nz, nzA);
    int s4 = cert_sign(2 * V * V - nv - nw - nz, 2 * VA * VA + nvA + nwA + nzA);
    if (s1 < 0 || s2 < 0 || s3 < 0 || s4 < 0) return -1;
    if (s1 == 0 || s2 == 0 || s3 == 0 || s4 == 0) return 0;
    return 1;
  }
  __host__ __device__ int side(const V3 &p) const {
    double Q = V * dot(p, p) - dot(p, N);
    double QA = VA * dot(p, p) + dot(vabs(p), NA);
    int s = cert_sign(Q, QA);
    return V > 0 ? s : -s;
  }
  __host__ __device__ double level() const { return dot(N, N) / (4.0 * V * V); }
  __host__ __device__ double level_err() const {
    double nn = dot(N, N), nnA = dot(NA, NA);
    return 8.0 * kCert * (1.0 + nnA / nn + (VA * VA) / (V * V));
  }
};

}  // namespace mhgp
