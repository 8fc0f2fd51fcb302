/* Oracle-only arbitrary-precision integers (test infrastructure, never shipped). */
#include "bigint.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void bi_die(const char *what) {
  fprintf(stderr, "oracle bigint: %s\n", what);
  abort();
}

static void bi_norm(bi *a) {
  while (a->n > 0 && a->d[a->n - 1] == 0) a->n--;
  if (a->n == 0) a->s = 0;
}

void bi_zero(bi *a) { a->n = 0; a->s = 0; }

void bi_set_u64(bi *a, uint64_t v) {
  a->d[0] = (uint32_t)v;
  a->d[1] = (uint32_t)(v >> 32);
  a->n = 2;
  a->s = 1;
  bi_norm(a);
}

void bi_set_i64(bi *a, int64_t v) {
  if (v < 0) {
    bi_set_u64(a, (uint64_t)0 - (uint64_t)v);
    if (a->n) a->s = -1;
  } else {
    bi_set_u64(a, (uint64_t)v);
  }
}

void bi_copy(bi *r, const bi *a) {
  if (r == a) return;
  r->n = a->n;
  r->s = a->s;
  memcpy(r->d, a->d, sizeof(uint32_t) * (size_t)a->n);
}

int bi_cmpabs(const bi *a, const bi *b) {
  if (a->n != b->n) return a->n < b->n ? -1 : 1;
  for (int i = a->n - 1; i >= 0; --i)
    if (a->d[i] != b->d[i]) return a->d[i] < b->d[i] ? -1 : 1;
  return 0;
}

/* |a| + |b| into r (r may alias) */
static void addabs(bi *r, const bi *a, const bi *b) {
  const bi *x = a->n >= b->n ? a : b;
  const bi *y = a->n >= b->n ? b : a;
  bi t;
  uint64_t c = 0;
  int i;
  for (i = 0; i < y->n; ++i) {
    c += (uint64_t)x->d[i] + y->d[i];
    t.d[i] = (uint32_t)c;
    c >>= 32;
  }
  for (; i < x->n; ++i) {
    c += x->d[i];
    t.d[i] = (uint32_t)c;
    c >>= 32;
  }
  t.n = x->n;
  if (c) {
    if (t.n >= BI_LIMBS) bi_die("capacity exceeded in add");
    t.d[t.n++] = (uint32_t)c;
  }
  t.s = 1;
  bi_norm(&t);
  bi_copy(r, &t);
}

/* |a| - |b| into r, requires |a| >= |b| */
static void subabs(bi *r, const bi *a, const bi *b) {
  bi t;
  int64_t br = 0;
  int i;
  for (i = 0; i < b->n; ++i) {
    int64_t v = (int64_t)a->d[i] - b->d[i] - br;
    br = v < 0;
    t.d[i] = (uint32_t)(v + (br ? ((int64_t)1 << 32) : 0));
  }
  for (; i < a->n; ++i) {
    int64_t v = (int64_t)a->d[i] - br;
    br = v < 0;
    t.d[i] = (uint32_t)(v + (br ? ((int64_t)1 << 32) : 0));
  }
  if (br) bi_die("subabs underflow");
  t.n = a->n;
  t.s = 1;
  bi_norm(&t);
  bi_copy(r, &t);
}

static void signed_add(bi *r, const bi *a, int sa, const bi *b, int sb) {
  if (sa == 0) { bi_copy(r, b); if (r->n) r->s = sb; return; }
  if (sb == 0) { bi_copy(r, a); if (r->n) r->s = sa; return; }
  if (sa == sb) {
    addabs(r, a, b);
    if (r->n) r->s = sa;
    return;
  }
  int c = bi_cmpabs(a, b);
  if (c == 0) { bi_zero(r); return; }
  if (c > 0) { subabs(r, a, b); if (r->n) r->s = sa; }
  else { subabs(r, b, a); if (r->n) r->s = sb; }
}

void bi_add(bi *r, const bi *a, const bi *b) { signed_add(r, a, a->s, b, b->s); }
void bi_sub(bi *r, const bi *a, const bi *b) { signed_add(r, a, a->s, b, -b->s); }

void bi_neg(bi *r, const bi *a) { bi_copy(r, a); r->s = -r->s; }

void bi_mul(bi *r, const bi *a, const bi *b) {
  if (a->s == 0 || b->s == 0) { bi_zero(r); return; }
  if (a->n + b->n > BI_LIMBS) bi_die("capacity exceeded in mul");
  bi t;
  memset(t.d, 0, sizeof(uint32_t) * (size_t)(a->n + b->n));
  for (int i = 0; i < a->n; ++i) {
    uint64_t c = 0;
    uint64_t ai = a->d[i];
    for (int j = 0; j < b->n; ++j) {
      c += ai * b->d[j] + t.d[i + j];
      t.d[i + j] = (uint32_t)c;
      c >>= 32;
    }
    t.d[i + b->n] = (uint32_t)c;
  }
  t.n = a->n + b->n;
  t.s = a->s * b->s;
  bi_norm(&t);
  bi_copy(r, &t);
}

void bi_mul_i64(bi *r, const bi *a, int64_t m) {
  bi t;
  bi_set_i64(&t, m);
  bi_mul(r, a, &t);
}

void bi_shl(bi *r, const bi *a, int k) {
  if (a->s == 0) { bi_zero(r); return; }
  int w = k / 32, s = k % 32;
  if (a->n + w + 1 > BI_LIMBS) bi_die("capacity exceeded in shl");
  bi t;
  memset(t.d, 0, sizeof(uint32_t) * (size_t)(a->n + w + 1));
  for (int i = 0; i < a->n; ++i) {
    uint64_t v = (uint64_t)a->d[i] << s;
    t.d[i + w] |= (uint32_t)v;
    t.d[i + w + 1] |= (uint32_t)(v >> 32);
  }
  t.n = a->n + w + 1;
  t.s = a->s;
  bi_norm(&t);
  bi_copy(r, &t);
}

static void shr_abs(bi *r, const bi *a, int k) {
  int w = k / 32, s = k % 32;
  if (w >= a->n) { bi_zero(r); return; }
  bi t;
  t.n = a->n - w;
  for (int i = 0; i < t.n; ++i) {
    uint64_t lo = a->d[i + w];
    uint64_t hi = (i + w + 1 < a->n) ? a->d[i + w + 1] : 0;
    t.d[i] = (uint32_t)(((hi << 32) | lo) >> s);
  }
  t.s = a->s;
  bi_norm(&t);
  bi_copy(r, &t);
}

int bi_cmp(const bi *a, const bi *b) {
  if (a->s != b->s) return a->s < b->s ? -1 : 1;
  int c = bi_cmpabs(a, b);
  return a->s >= 0 ? c : -c;
}

int bi_bitlen(const bi *a) {
  if (a->n == 0) return 0;
  uint32_t top = a->d[a->n - 1];
  int b = 0;
  while (top) { b++; top >>= 1; }
  return (a->n - 1) * 32 + b;
}

static int ctz_abs(const bi *a) {
  for (int i = 0; i < a->n; ++i)
    if (a->d[i]) {
      int b = 0;
      uint32_t v = a->d[i];
      while (!(v & 1)) { v >>= 1; b++; }
      return i * 32 + b;
    }
  return 0;
}

void bi_gcd(bi *r, const bi *a, const bi *b) {
  bi x, y;
  bi_copy(&x, a); if (x.s) x.s = 1;
  bi_copy(&y, b); if (y.s) y.s = 1;
  if (x.s == 0) { bi_copy(r, &y); return; }
  if (y.s == 0) { bi_copy(r, &x); return; }
  int zx = ctz_abs(&x), zy = ctz_abs(&y);
  int k = zx < zy ? zx : zy;
  shr_abs(&x, &x, zx);
  shr_abs(&y, &y, zy);
  while (y.s != 0) {
    int c = bi_cmpabs(&x, &y);
    if (c > 0) { bi t = x; x = y; y = t; }
    subabs(&y, &y, &x);
    if (y.s) shr_abs(&y, &y, ctz_abs(&y));
  }
  bi_shl(r, &x, k);
}

void bi_divmod(bi *q, bi *rem, const bi *a, const bi *b) {
  if (b->s <= 0 || a->s < 0) bi_die("divmod requires a >= 0 and b > 0");
  bi quo, r, one;
  bi_zero(&quo);
  bi_zero(&r);
  bi_set_u64(&one, 1);
  /* restoring binary long division, one quotient bit per step */
  for (int i = bi_bitlen(a) - 1; i >= 0; --i) {
    bi_shl(&r, &r, 1);
    if ((a->d[i / 32] >> (i % 32)) & 1u) bi_add(&r, &r, &one);
    bi_shl(&quo, &quo, 1);
    if (bi_cmpabs(&r, b) >= 0) {
      subabs(&r, &r, b);
      bi_add(&quo, &quo, &one);
    }
  }
  if (q) bi_copy(q, &quo);
  if (rem) bi_copy(rem, &r);
}

static uint32_t divsmall(bi *a, uint32_t m) {
  uint64_t rem = 0;
  for (int i = a->n - 1; i >= 0; --i) {
    uint64_t cur = (rem << 32) | a->d[i];
    a->d[i] = (uint32_t)(cur / m);
    rem = cur % m;
  }
  int s = a->s;
  bi_norm(a);
  if (a->n) a->s = s;
  return (uint32_t)rem;
}

int bi_to_dec(const bi *a, char *buf, int cap) {
  if (a->s == 0) {
    if (cap < 2) return -1;
    buf[0] = '0'; buf[1] = 0;
    return 1;
  }
  bi t;
  bi_copy(&t, a);
  t.s = 1;
  char tmp[8192];
  int len = 0;
  while (t.n) {
    uint32_t r = divsmall(&t, 1000000000u);
    for (int i = 0; i < 9; ++i) {
      tmp[len++] = (char)('0' + r % 10);
      r /= 10;
      if (len >= (int)sizeof(tmp) - 2) bi_die("decimal buffer");
    }
  }
  while (len > 1 && tmp[len - 1] == '0') len--;
  int pos = 0;
  if (a->s < 0) buf[pos++] = '-';
  if (pos + len + 1 > cap) return -1;
  for (int i = len - 1; i >= 0; --i) buf[pos++] = tmp[i];
  buf[pos] = 0;
  return pos;
}

double bi_to_double(const bi *a) {
  double v = 0.0;
  for (int i = a->n - 1; i >= 0; --i) v = v * 4294967296.0 + (double)a->d[i];
  return a->s < 0 ? -v : v;
}
