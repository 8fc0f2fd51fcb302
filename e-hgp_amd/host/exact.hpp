// Host multiprecision fallback of the MI355X backend.
//
// The GPU decides almost every predicate with certified FP64 filters; what
// they leave undecided is settled here exactly.  Coordinates are binary64,
// hence dyadic: the cloud is rescaled once to integers on a common exponent
// and every predicate becomes the sign of an integer polynomial, evaluated
// with the small arbitrary-precision integer below (the role Boost's cpp_int
// plays in the reference, morsehgp3d/include/morsehgp3d/exact/integer.hpp:13).
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <vector>

namespace mhgp {

class BigInt {
 public:
  BigInt() = default;
  explicit BigInt(int64_t v);
  static BigInt from_dyadic(int64_t significand, int shift);  // significand << shift
  int sign() const { return sign_; }
  BigInt operator+(const BigInt &o) const;
  BigInt operator-(const BigInt &o) const;
  BigInt operator*(const BigInt &o) const;
  BigInt operator-() const;
  friend int compare(const BigInt &a, const BigInt &b);
  double to_double() const;
  // Leading 64 bits as a double and the count of bits below them:
  // value = top * 2^shift up to a relative 2^-52 (no overflow at any size).
  double top_bits(int &shift) const;

 private:
  static int cmp_mag(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b);
  static std::vector<uint32_t> add_mag(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b);
  static std::vector<uint32_t> sub_mag(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b);
  void trim();
  std::vector<uint32_t> mag_;  // little endian
  int sign_ = 0;
};

// Exact cloud in canonical PointId order, coordinates = integer * 2^e0.
// The integer coordinates of a point are made on its first use (most exact
// work touches few points); coord() may be called from several threads.
class ExactCloud {
 public:
  ExactCloud() = default;
  void load(const double *xyz, int64_t n);  // xyz interleaved, canonical order; kept by reference
  bool loaded() const { return n_ > 0; }
  int64_t size() const { return n_; }
  const BigInt &coord(int64_t i, int a) const {
    if (state_[i].load(std::memory_order_acquire) != 2) make(i);
    return P_[3 * i + a];
  }
  BigInt sq_dist(int64_t i, int64_t j) const;
  // sign(|q-i|^2 - |q-j|^2)
  int cmp_dist(int64_t q, int64_t i, int64_t j) const;
  int e0() const { return e0_; }

 private:
  void make(int64_t i) const;
  int64_t n_ = 0;
  int e0_ = 0;
  const double *xyz_ = nullptr;
  mutable std::vector<BigInt> P_;
  mutable std::unique_ptr<std::atomic<uint8_t>[]> state_;  // per point: 0 none, 1 in progress, 2 made
};

// Exact support analysis: circumsphere with centre in the affine hull of
// 1..4 points, well-centring by barycentric signs, side of a point and the
// squared radius as a ratio num/den (den > 0) in scaled integer units.
struct ExactSupport {
  int m = 0;
  int64_t ids[4] = {0, 0, 0, 0};
  bool dependent = false;
  bool well = false;
  BigInt a[3], b[3], d[3];
  BigInt D, al, be;  // size 3
  BigInt V, N[3];    // size 4
  BigInt lnum, lden;

  bool analyze(const ExactCloud &c, const int64_t *ids, int m);
  int side(const ExactCloud &c, int64_t x) const;  // -1 inside, 0 on, +1 outside
};

// sign(l1 - l2) for squared radii given as num/den with den > 0
int compare_levels(const BigInt &n1, const BigInt &d1, const BigInt &n2, const BigInt &d2);

// num / den * 2^exp2 in binary64, relative error below 1e-15 (num, den > 0).
double ratio_to_double(const BigInt &num, const BigInt &den, int exp2);

}  // namespace mhgp
