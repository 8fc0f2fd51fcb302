// Host multiprecision fallback of the MI355X backend (see exact.hpp).
#include "exact.hpp"

#include <thread>

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace mhgp {

BigInt::BigInt(int64_t v) {
  if (v == 0) return;
  uint64_t m = v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v;
  sign_ = v < 0 ? -1 : 1;
  mag_.push_back((uint32_t)m);
  if (m >> 32) mag_.push_back((uint32_t)(m >> 32));
}

BigInt BigInt::from_dyadic(int64_t significand, int shift) {
  BigInt r(significand);
  if (r.sign_ == 0 || shift == 0) return r;
  if (shift < 0) throw std::logic_error("negative dyadic shift");
  const int w = shift / 32, s = shift % 32;
  std::vector<uint32_t> out(r.mag_.size() + w + 1, 0);
  for (size_t i = 0; i < r.mag_.size(); ++i) {
    uint64_t v = (uint64_t)r.mag_[i] << s;
    out[i + w] |= (uint32_t)v;
    out[i + w + 1] |= (uint32_t)(v >> 32);
  }
  r.mag_ = std::move(out);
  r.trim();
  return r;
}

void BigInt::trim() {
  while (!mag_.empty() && mag_.back() == 0) mag_.pop_back();
  if (mag_.empty()) sign_ = 0;
}

int BigInt::cmp_mag(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
  if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
  for (size_t i = a.size(); i-- > 0;)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

std::vector<uint32_t> BigInt::add_mag(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
  const std::vector<uint32_t> &x = a.size() >= b.size() ? a : b;
  const std::vector<uint32_t> &y = a.size() >= b.size() ? b : a;
  std::vector<uint32_t> r(x.size() + 1, 0);
  uint64_t c = 0;
  for (size_t i = 0; i < x.size(); ++i) {
    c += (uint64_t)x[i] + (i < y.size() ? y[i] : 0);
    r[i] = (uint32_t)c;
    c >>= 32;
  }
  r[x.size()] = (uint32_t)c;
  return r;
}

// |a| - |b| with |a| >= |b|
std::vector<uint32_t> BigInt::sub_mag(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
  std::vector<uint32_t> r(a.size(), 0);
  int64_t br = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    int64_t v = (int64_t)a[i] - (i < b.size() ? b[i] : 0) - br;
    br = v < 0;
    r[i] = (uint32_t)(v + (br ? ((int64_t)1 << 32) : 0));
  }
  return r;
}

BigInt BigInt::operator+(const BigInt &o) const {
  if (sign_ == 0) return o;
  if (o.sign_ == 0) return *this;
  BigInt r;
  if (sign_ == o.sign_) {
    r.mag_ = add_mag(mag_, o.mag_);
    r.sign_ = sign_;
  } else {
    int c = cmp_mag(mag_, o.mag_);
    if (c == 0) return BigInt();
    r.mag_ = c > 0 ? sub_mag(mag_, o.mag_) : sub_mag(o.mag_, mag_);
    r.sign_ = c > 0 ? sign_ : o.sign_;
  }
  r.trim();
  return r;
}

BigInt BigInt::operator-() const {
  BigInt r = *this;
  r.sign_ = -r.sign_;
  return r;
}

BigInt BigInt::operator-(const BigInt &o) const { return *this + (-o); }

BigInt BigInt::operator*(const BigInt &o) const {
  if (sign_ == 0 || o.sign_ == 0) return BigInt();
  BigInt r;
  r.mag_.assign(mag_.size() + o.mag_.size(), 0);
  for (size_t i = 0; i < mag_.size(); ++i) {
    uint64_t c = 0;
    const uint64_t ai = mag_[i];
    for (size_t j = 0; j < o.mag_.size(); ++j) {
      c += ai * o.mag_[j] + r.mag_[i + j];
      r.mag_[i + j] = (uint32_t)c;
      c >>= 32;
    }
    r.mag_[i + o.mag_.size()] = (uint32_t)c;
  }
  r.sign_ = sign_ * o.sign_;
  r.trim();
  return r;
}

int compare(const BigInt &a, const BigInt &b) {
  if (a.sign_ != b.sign_) return a.sign_ < b.sign_ ? -1 : 1;
  int c = BigInt::cmp_mag(a.mag_, b.mag_);
  return a.sign_ >= 0 ? c : -c;
}

double BigInt::to_double() const {
  double v = 0.0;
  for (size_t i = mag_.size(); i-- > 0;) v = v * 4294967296.0 + (double)mag_[i];
  return sign_ < 0 ? -v : v;
}

double BigInt::top_bits(int &shift) const {
  shift = 0;
  if (mag_.empty()) return 0.0;
  const size_t w = mag_.size();
  uint64_t top = mag_[w - 1];
  if (w >= 2) top = (top << 32) | mag_[w - 2];
  int lead = 64 - __builtin_clzll(top);  // bits of top in use
  // fill up to 64 leading bits from the next word
  if (w >= 3 && lead < 64) {
    const int need = 64 - lead;
    top = (top << need) | ((uint64_t)mag_[w - 3] >> (32 - need));
    shift = (int)(32 * (w - 3)) + (32 - need);
  } else {
    shift = w >= 2 ? (int)(32 * (w - 2)) : 0;
  }
  const double v = (double)top;
  return sign_ < 0 ? -v : v;
}

double ratio_to_double(const BigInt &num, const BigInt &den, int exp2) {
  int sn = 0, sd = 0;
  const double tn = num.top_bits(sn), td = den.top_bits(sd);
  return std::ldexp(tn / td, sn - sd + exp2);
}

static void decode_binary64(double x, int64_t &sig, int &exp) {
  uint64_t b;
  std::memcpy(&b, &x, 8);
  uint64_t e = (b >> 52) & 0x7ff, f = b & ((1ull << 52) - 1);
  if (e == 0x7ff) throw std::invalid_argument("non-finite coordinate");
  uint64_t mag = f;
  int ex = -1074;
  if (e) { mag |= 1ull << 52; ex = (int)e - 1075; }
  sig = (b >> 63) ? -(int64_t)mag : (int64_t)mag;
  exp = ex;
  if (sig == 0) { exp = 0; return; }
  while ((sig & 1) == 0) { sig /= 2; exp++; }
}

void ExactCloud::load(const double *xyz, int64_t n) {
  n_ = n;
  int emin = 0;
  bool any = false;
  for (int64_t i = 0; i < 3 * n; ++i) {
    int64_t s; int e;
    decode_binary64(xyz[i], s, e);
    if (s == 0) continue;
    if (!any || e < emin) emin = e;
    any = true;
  }
  e0_ = emin;
  xyz_ = xyz;
  P_.assign(3 * n, BigInt());
  state_.reset(new std::atomic<uint8_t>[n]);
  for (int64_t i = 0; i < n; ++i) state_[i].store(0, std::memory_order_relaxed);
}

void ExactCloud::make(int64_t i) const {
  uint8_t expect = 0;
  if (state_[i].compare_exchange_strong(expect, 1, std::memory_order_acq_rel)) {
    for (int a = 0; a < 3; ++a) {
      int64_t s; int e;
      decode_binary64(xyz_[3 * i + a], s, e);
      P_[3 * i + a] = s ? BigInt::from_dyadic(s, e - e0_) : BigInt();
    }
    state_[i].store(2, std::memory_order_release);
    return;
  }
  while (state_[i].load(std::memory_order_acquire) != 2) std::this_thread::yield();
}

BigInt ExactCloud::sq_dist(int64_t i, int64_t j) const {
  BigInt s;
  for (int a = 0; a < 3; ++a) {
    BigInt d = coord(i, a) - coord(j, a);
    s = s + d * d;
  }
  return s;
}

int ExactCloud::cmp_dist(int64_t q, int64_t i, int64_t j) const {
  return compare(sq_dist(q, i), sq_dist(q, j));
}

static BigInt dot3(const BigInt *x, const BigInt *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; }
static void cross3(const BigInt *x, const BigInt *y, BigInt *r) {
  r[0] = x[1] * y[2] - x[2] * y[1];
  r[1] = x[2] * y[0] - x[0] * y[2];
  r[2] = x[0] * y[1] - x[1] * y[0];
}

bool ExactSupport::analyze(const ExactCloud &c, const int64_t *id, int mm) {
  m = mm;
  for (int i = 0; i < m; ++i) ids[i] = id[i];
  dependent = false;
  well = true;
  if (m == 1) {
    lnum = BigInt();
    lden = BigInt(1);
    return true;
  }
  for (int a = 0; a < 3; ++a) this->a[a] = c.coord(id[1], a) - c.coord(id[0], a);
  if (m == 2) {
    lnum = dot3(this->a, this->a);
    lden = BigInt(4);
    return true;
  }
  for (int a = 0; a < 3; ++a) b[a] = c.coord(id[2], a) - c.coord(id[0], a);
  if (m == 3) {
    BigInt aa = dot3(this->a, this->a), bb = dot3(b, b), ab = dot3(this->a, b);
    D = aa * bb - ab * ab;
    if (D.sign() == 0) { dependent = true; return false; }
    al = bb * (aa - ab);
    be = aa * (bb - ab);
    BigInt rest = D + D - al - be;
    well = al.sign() > 0 && be.sign() > 0 && rest.sign() > 0;
    BigInt amb[3] = {this->a[0] - b[0], this->a[1] - b[1], this->a[2] - b[2]};
    lnum = aa * bb * dot3(amb, amb);
    lden = D * BigInt(4);
    return true;
  }
  for (int a = 0; a < 3; ++a) d[a] = c.coord(id[3], a) - c.coord(id[0], a);
  BigInt bxd[3], dxa[3], axb[3];
  cross3(b, d, bxd);
  cross3(d, this->a, dxa);
  cross3(this->a, b, axb);
  V = dot3(this->a, bxd);
  if (V.sign() == 0) { dependent = true; return false; }
  BigInt aa = dot3(this->a, this->a), bb = dot3(b, b), dd = dot3(d, d);
  for (int a = 0; a < 3; ++a) N[a] = aa * bxd[a] + bb * dxa[a] + dd * axb[a];
  BigInt nv = dot3(N, bxd), nw = dot3(N, dxa), nz = dot3(N, axb);
  BigInt v2 = V * V;
  BigInt nu = v2 + v2 - nv - nw - nz;
  well = nv.sign() > 0 && nw.sign() > 0 && nz.sign() > 0 && nu.sign() > 0;
  lnum = dot3(N, N);
  lden = v2 * BigInt(4);
  return true;
}

int ExactSupport::side(const ExactCloud &c, int64_t x) const {
  BigInt p[3];
  for (int a = 0; a < 3; ++a) p[a] = c.coord(x, a) - c.coord(ids[0], a);
  if (m == 1) return (p[0].sign() || p[1].sign() || p[2].sign()) ? 1 : 0;
  if (m == 2) {
    BigInt q[3];
    for (int a = 0; a < 3; ++a) q[a] = c.coord(x, a) - c.coord(ids[1], a);
    return dot3(p, q).sign();
  }
  if (m == 3) {
    BigInt t = D * dot3(p, p) - al * dot3(p, a) - be * dot3(p, b);
    return t.sign();
  }
  BigInt t = V * dot3(p, p) - dot3(p, N);
  return t.sign() * V.sign();
}

int compare_levels(const BigInt &n1, const BigInt &d1, const BigInt &n2, const BigInt &d2) {
  return compare(n1 * d2, n2 * d1);
}

}  // namespace mhgp
