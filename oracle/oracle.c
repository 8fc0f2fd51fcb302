/* CPU oracle for the MorseHGP3D hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product never links or calls it.
 *
 * Every decision here is exact: binary64 inputs are decoded to dyadic
 * integers on one common exponent and all predicates are evaluated with the
 * bigint type of bigint.c.  Each entry point cites the reference function whose
 * observable result it restates.
 */
#include "bigint.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_OK 0
#define OR_DUPLICATE -1
#define OR_NONFINITE -2
#define OR_RANGE -3
#define OR_CAPACITY -4
#define OR_ARGUMENT -5

/* ------------------------------------------------------------------------ */
/* binary64 helpers                                                          */
/* ------------------------------------------------------------------------ */

static uint64_t bits_of(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return b;
}

/* -0 becomes +0 (reference: exact/binary64.hpp:18 canonicalize_binary64_bits) */
static uint64_t canon_bits(uint64_t b) { return b == 0x8000000000000000ULL ? 0 : b; }

/* monotone unsigned key (reference: exact/binary64.hpp:25-31) */
static uint64_t order_key(uint64_t b) {
  b = canon_bits(b);
  return (b >> 63) ? ~b : (b ^ 0x8000000000000000ULL);
}

static int decode(uint64_t bits, int64_t *sig, int *exp) {
  uint64_t e = (bits >> 52) & 0x7ff, f = bits & ((1ULL << 52) - 1);
  if (e == 0x7ff) return OR_NONFINITE;
  uint64_t mag = f;
  int ex = -1074;
  if (e) { mag |= 1ULL << 52; ex = (int)e - 1075; }
  *sig = (bits >> 63) ? -(int64_t)mag : (int64_t)mag;
  *exp = ex;
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* exact cloud: coordinates as integers times 2^e0                           */
/* ------------------------------------------------------------------------ */

typedef struct { bi c[3]; } ipt;

typedef struct {
  int64_t n;
  int e0;
  ipt *P;
  const double *X;
} ecloud;

static int ecloud_init(ecloud *cl, const double *xyz, int64_t n) {
  cl->n = n;
  cl->X = xyz;
  int emin = 0, emax = 0, any = 0;
  for (int64_t i = 0; i < 3 * n; ++i) {
    int64_t s; int e;
    if (decode(canon_bits(bits_of(xyz[i])), &s, &e)) return OR_NONFINITE;
    if (s == 0) continue;
    /* strip trailing zeros so the exponent is the true LSB */
    while ((s & 1) == 0) { s /= 2; e++; }
    int top = e + 64 - __builtin_clzll((uint64_t)(s < 0 ? -s : s));
    if (!any) { emin = e; emax = top; any = 1; }
    if (e < emin) emin = e;
    if (top > emax) emax = top;
  }
  if (any && emax - emin > 500) return OR_RANGE;
  cl->e0 = emin;
  cl->P = (ipt *)malloc(sizeof(ipt) * (size_t)(n ? n : 1));
  if (!cl->P) return OR_CAPACITY;
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      int64_t s = 0; int e = 0;
      decode(canon_bits(bits_of(xyz[3 * i + a])), &s, &e);
      while (s != 0 && (s & 1) == 0) { s /= 2; e++; }
      bi_set_i64(&cl->P[i].c[a], s);
      if (s) bi_shl(&cl->P[i].c[a], &cl->P[i].c[a], e - emin);
    }
  return OR_OK;
}

static void ecloud_free(ecloud *cl) { free(cl->P); cl->P = NULL; }

/* ------------------------------------------------------------------------ */
/* canonical PointIds                                                        */
/* reference: morsehgp3d/src/cpu/spatial/point_cloud.cpp:40-95               */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t k[3]; int64_t src; } ckey;

static int ckey_cmp(const void *a, const void *b) {
  const ckey *x = a, *y = b;
  for (int i = 0; i < 3; ++i)
    if (x->k[i] != y->k[i]) return x->k[i] < y->k[i] ? -1 : 1;
  return x->src < y->src ? -1 : (x->src > y->src);
}

/* perm[id] = source index of the point whose canonical PointId is id. */
int or_canonical_order(const double *xyz, int64_t n, int64_t *perm) {
  if (n <= 0) return OR_ARGUMENT;
  ckey *keys = malloc(sizeof(ckey) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      uint64_t b = bits_of(xyz[3 * i + a]);
      if (((b >> 52) & 0x7ff) == 0x7ff) { free(keys); return OR_NONFINITE; }
      keys[i].k[a] = order_key(b);
    }
    keys[i].src = i;
  }
  qsort(keys, (size_t)n, sizeof(ckey), ckey_cmp);
  for (int64_t i = 1; i < n; ++i)
    if (!memcmp(keys[i - 1].k, keys[i].k, sizeof(keys[i].k))) { free(keys); return OR_DUPLICATE; }
  for (int64_t i = 0; i < n; ++i) perm[i] = keys[i].src;
  free(keys);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* exact Morton LBVH snapshot                                                */
/* reference: morsehgp3d/src/cpu/spatial/lbvh.cpp:46-151 (witnesses, frame,  */
/* interleave), 237-298 (exact quantisation), 484-560 (build), 831-906       */
/* (find_split / build_range, strict postorder)                             */
/* ------------------------------------------------------------------------ */

#define MORTON_BITS 21

static uint64_t interleave(const uint64_t q[3]) {
  uint64_t code = 0;
  for (int b = 0; b < MORTON_BITS; ++b) {
    code |= ((q[0] >> b) & 1ULL) << (3 * b + 2);
    code |= ((q[1] >> b) & 1ULL) << (3 * b + 1);
    code |= ((q[2] >> b) & 1ULL) << (3 * b);
  }
  return code;
}

/* floor(2^21 (x - l) / (u - l)) on exact integers */
static uint64_t quantize(const bi *x, const bi *l, const bi *u) {
  bi ext, del, num, q;
  bi_sub(&ext, u, l);
  if (ext.s == 0) return 0;
  if (bi_cmp(x, u) == 0) return (1ULL << MORTON_BITS) - 1;
  bi_sub(&del, x, l);
  bi_shl(&num, &del, MORTON_BITS);
  bi_divmod(&q, NULL, &num, &ext);
  uint64_t v = 0;
  for (int i = q.n - 1; i >= 0; --i) v = (v << 32) | q.d[i];
  return v;
}

typedef struct { uint64_t code; uint64_t id; } leafrec;

static int leaf_cmp(const void *a, const void *b) {
  const leafrec *x = a, *y = b;
  if (x->code != y->code) return x->code < y->code ? -1 : 1;
  return x->id < y->id ? -1 : (x->id > y->id);
}

typedef struct {
  uint64_t lo[3], hi[3], left, right, begin, end;
} noderec;

typedef struct {
  const double *X;
  leafrec *leaves;
  noderec *nodes;
  int64_t count;
  uint64_t maxdepth;
} lbvh_builder;

static uint64_t key_at(const double *X, uint64_t id, int a) { return order_key(bits_of(X[3 * id + a])); }

static uint64_t lower_wit(const double *X, uint64_t l, uint64_t r, int a) {
  uint64_t kl = key_at(X, l, a), kr = key_at(X, r, a);
  return (kr < kl || (kr == kl && r < l)) ? r : l;
}
static uint64_t upper_wit(const double *X, uint64_t l, uint64_t r, int a) {
  uint64_t kl = key_at(X, l, a), kr = key_at(X, r, a);
  return (kr > kl || (kr == kl && r < l)) ? r : l;
}

static uint64_t find_split(const leafrec *L, uint64_t b, uint64_t e) {
  uint64_t fc = L[b].code, lc = L[e - 1].code;
  if (fc == lc) return b + (e - b) / 2;
  uint64_t diff = fc ^ lc;
  int hb = 63 - __builtin_clzll(diff);
  uint64_t mask = 1ULL << hb;
  uint64_t lo = b + 1, hi = e;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if ((L[mid].code & mask) == 0) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static uint64_t build_range(lbvh_builder *B, uint64_t b, uint64_t e, uint64_t depth) {
  if (depth > B->maxdepth) B->maxdepth = depth;
  noderec nd;
  if (e - b == 1) {
    for (int a = 0; a < 3; ++a) nd.lo[a] = nd.hi[a] = B->leaves[b].id;
    nd.left = nd.right = ~0ULL;
    nd.begin = b; nd.end = e;
    B->nodes[B->count] = nd;
    return (uint64_t)B->count++;
  }
  uint64_t s = find_split(B->leaves, b, e);
  uint64_t l = build_range(B, b, s, depth + 1);
  uint64_t r = build_range(B, s, e, depth + 1);
  for (int a = 0; a < 3; ++a) {
    nd.lo[a] = lower_wit(B->X, B->nodes[l].lo[a], B->nodes[r].lo[a], a);
    nd.hi[a] = upper_wit(B->X, B->nodes[l].hi[a], B->nodes[r].hi[a], a);
  }
  nd.left = l; nd.right = r; nd.begin = b; nd.end = e;
  B->nodes[B->count] = nd;
  return (uint64_t)B->count++;
}

/* xyz must already be in canonical PointId order.
 * leaves: 2n u64 (code, id); nodes: 10 u64 per node, 2n-1 nodes (postorder);
 * counters: point_count, node_count, maximum_depth, collision groups, max collision. */
int or_lbvh_snapshot(const double *xyz, int64_t n, uint64_t *leaves_out, uint64_t *nodes_out,
                     uint64_t *counters, uint64_t *codes_by_id) {
  if (n <= 0) return OR_ARGUMENT;
  ecloud cl;
  int rc = ecloud_init(&cl, xyz, n);
  if (rc) return rc;
  int lo_id[3] = {0, 0, 0}, hi_id[3] = {0, 0, 0};
  for (int64_t i = 1; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      if (key_at(xyz, (uint64_t)i, a) < key_at(xyz, (uint64_t)lo_id[a], a)) lo_id[a] = (int)i;
      if (key_at(xyz, (uint64_t)i, a) > key_at(xyz, (uint64_t)hi_id[a], a)) hi_id[a] = (int)i;
    }
  leafrec *L = malloc(sizeof(leafrec) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    uint64_t q[3];
    for (int a = 0; a < 3; ++a)
      q[a] = quantize(&cl.P[i].c[a], &cl.P[lo_id[a]].c[a], &cl.P[hi_id[a]].c[a]);
    L[i].code = interleave(q);
    L[i].id = (uint64_t)i;
    if (codes_by_id) codes_by_id[i] = L[i].code;
  }
  qsort(L, (size_t)n, sizeof(leafrec), leaf_cmp);
  uint64_t groups = 0, maxg = 0;
  for (int64_t b = 0; b < n;) {
    int64_t e = b + 1;
    while (e < n && L[e].code == L[b].code) e++;
    if (e - b > 1) { groups++; if ((uint64_t)(e - b) > maxg) maxg = (uint64_t)(e - b); }
    b = e;
  }
  lbvh_builder B = {xyz, L, malloc(sizeof(noderec) * (size_t)(2 * n - 1)), 0, 0};
  build_range(&B, 0, (uint64_t)n, 0);
  for (int64_t i = 0; i < n; ++i) { leaves_out[2 * i] = L[i].code; leaves_out[2 * i + 1] = L[i].id; }
  for (int64_t i = 0; i < 2 * n - 1; ++i) {
    noderec *nd = &B.nodes[i];
    uint64_t *o = nodes_out + 10 * i;
    for (int a = 0; a < 3; ++a) { o[a] = nd->lo[a]; o[3 + a] = nd->hi[a]; }
    o[6] = nd->left; o[7] = nd->right; o[8] = nd->begin; o[9] = nd->end;
  }
  counters[0] = (uint64_t)n;
  counters[1] = (uint64_t)(2 * n - 1);
  counters[2] = B.maxdepth;
  counters[3] = groups;
  counters[4] = maxg;
  free(B.nodes);
  free(L);
  ecloud_free(&cl);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* exact squared distances                                                   */
/* ------------------------------------------------------------------------ */

static void sq_dist(bi *r, const ipt *p, const ipt *q) {
  bi t, s;
  bi_zero(r);
  for (int a = 0; a < 3; ++a) {
    bi_sub(&t, &p->c[a], &q->c[a]);
    bi_mul(&s, &t, &t);
    bi_add(r, r, &s);
  }
}

/* fast float estimate with a certified relative error; returns -1/0/1 or 2 for undecided */
static int filt_cmp(double a, double b) {
  double d = a - b, m = fabs(a) + fabs(b);
  if (d > 1e-10 * m) return 1;
  if (-d > 1e-10 * m) return -1;
  return 2;
}

static double fdist2(const double *X, int64_t i, int64_t j) {
  double s = 0;
  for (int a = 0; a < 3; ++a) { double t = X[3 * i + a] - X[3 * j + a]; s += t * t; }
  return s;
}

/* ------------------------------------------------------------------------ */
/* brute-force k nearest neighbours (exact (d^2, id) order, self excluded)   */
/* reference semantics: morsehgp3d/src/cpu/spatial/brute_force.cpp (top-k by */
/* exact squared distance with PointId order inside an equal shell)          */
/* ------------------------------------------------------------------------ */

typedef struct { double f; int64_t id; } fcand;

static ecloud *g_knn_cl;
static int64_t g_knn_q;

static int exact_dist_cmp(const ecloud *cl, int64_t q, int64_t i, int64_t j) {
  bi di, dj;
  sq_dist(&di, &cl->P[q], &cl->P[i]);
  sq_dist(&dj, &cl->P[q], &cl->P[j]);
  return bi_cmp(&di, &dj);
}

static int knn_cmp(const void *a, const void *b) {
  const fcand *x = a, *y = b;
  int c = filt_cmp(x->f, y->f);
  if (c == 2) c = exact_dist_cmp(g_knn_cl, g_knn_q, x->id, y->id);
  if (c) return c;
  return x->id < y->id ? -1 : (x->id > y->id);
}

/* out: n*k ids; query = data point i, neighbours exclude i; also returns in
 * tie_flags[i] whether the (k+1)-th neighbour ties the k-th exactly. */
int or_knn(const double *xyz, int64_t n, int64_t k, int64_t *out, uint8_t *tie_flags) {
  if (k <= 0 || k >= n) return OR_ARGUMENT;
  ecloud cl;
  int rc = ecloud_init(&cl, xyz, n);
  if (rc) return rc;
  fcand *c = malloc(sizeof(fcand) * (size_t)n);
  g_knn_cl = &cl;
  for (int64_t q = 0; q < n; ++q) {
    int64_t m = 0;
    for (int64_t j = 0; j < n; ++j)
      if (j != q) { c[m].f = fdist2(xyz, q, j); c[m].id = j; m++; }
    g_knn_q = q;
    qsort(c, (size_t)m, sizeof(fcand), knn_cmp);
    for (int64_t t = 0; t < k; ++t) out[q * k + t] = c[t].id;
    if (tie_flags) tie_flags[q] = (k < m) && exact_dist_cmp(&cl, q, c[k - 1].id, c[k].id) == 0;
  }
  free(c);
  ecloud_free(&cl);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* exact EMST by Kruskal over the complete graph under kappa=(d^2,u,v)       */
/* reference: morsehgp3d/src/cpu/hierarchy/emst.cpp:326 and                  */
/* include/morsehgp3d/hierarchy/emst.hpp:24-30                               */
/* ------------------------------------------------------------------------ */

typedef struct { double f; int64_t u, v; } fedge;
static ecloud *g_emst_cl;

static int edge_cmp(const void *a, const void *b) {
  const fedge *x = a, *y = b;
  int c = filt_cmp(x->f, y->f);
  if (c == 2) {
    bi dx, dy;
    sq_dist(&dx, &g_emst_cl->P[x->u], &g_emst_cl->P[x->v]);
    sq_dist(&dy, &g_emst_cl->P[y->u], &g_emst_cl->P[y->v]);
    c = bi_cmp(&dx, &dy);
  }
  if (c) return c;
  if (x->u != y->u) return x->u < y->u ? -1 : 1;
  return x->v < y->v ? -1 : (x->v > y->v);
}

static int64_t uf_find(int64_t *p, int64_t x) {
  while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
  return x;
}

/* edges: 2(n-1) ids (u<v), in kappa order */
int or_emst(const double *xyz, int64_t n, int64_t *edges) {
  if (n < 1) return OR_ARGUMENT;
  if (n == 1) return OR_OK;
  ecloud cl;
  int rc = ecloud_init(&cl, xyz, n);
  if (rc) return rc;
  int64_t m = n * (n - 1) / 2;
  fedge *E = malloc(sizeof(fedge) * (size_t)m);
  if (!E) { ecloud_free(&cl); return OR_CAPACITY; }
  int64_t t = 0;
  for (int64_t u = 0; u < n; ++u)
    for (int64_t v = u + 1; v < n; ++v) { E[t].f = fdist2(xyz, u, v); E[t].u = u; E[t].v = v; t++; }
  g_emst_cl = &cl;
  qsort(E, (size_t)m, sizeof(fedge), edge_cmp);
  int64_t *par = malloc(sizeof(int64_t) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) par[i] = i;
  int64_t k = 0;
  for (int64_t i = 0; i < m && k < n - 1; ++i) {
    int64_t a = uf_find(par, E[i].u), b = uf_find(par, E[i].v);
    if (a == b) continue;
    par[a] = b;
    edges[2 * k] = E[i].u;
    edges[2 * k + 1] = E[i].v;
    k++;
  }
  free(par);
  free(E);
  ecloud_free(&cl);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* exact support analysis (1..4 points)                                      */
/* reference: morsehgp3d/include/morsehgp3d/exact/center.hpp:263-340          */
/* (circumcentre in the affine hull), support.hpp:41-135 (well-centring by   */
/* barycentric signs), support.hpp:814 (sphere side of a point)              */
/* ------------------------------------------------------------------------ */

typedef struct {
  int m;
  int64_t ids[4];
  int dependent;
  int well;
  bi a[3], b[3], d[3];
  bi D, al, be;  /* m = 3 */
  bi V, N[3];    /* m = 4 */
  bi lnum, lden; /* squared radius = lnum / lden in integer units (lden > 0) */
  bi cnum[3], cden; /* centre = cnum / cden in integer units (cden > 0) */
} sup_t;

static void vsub(bi r[3], const ipt *p, const ipt *q) {
  for (int a = 0; a < 3; ++a) bi_sub(&r[a], &p->c[a], &q->c[a]);
}
static void vdot(bi *r, const bi x[3], const bi y[3]) {
  bi t;
  bi_zero(r);
  for (int a = 0; a < 3; ++a) { bi_mul(&t, &x[a], &y[a]); bi_add(r, r, &t); }
}
static void vcross(bi r[3], const bi x[3], const bi y[3]) {
  bi t, u, res[3];
  for (int a = 0; a < 3; ++a) {
    int i = (a + 1) % 3, j = (a + 2) % 3;
    bi_mul(&t, &x[i], &y[j]);
    bi_mul(&u, &x[j], &y[i]);
    bi_sub(&res[a], &t, &u);
  }
  for (int a = 0; a < 3; ++a) bi_copy(&r[a], &res[a]);
}

/* Builds the circumball of the support in its affine hull.  Returns 0 when
 * the support is affinely dependent. */
static int support_analyze(sup_t *S, const ecloud *cl, const int64_t *ids, int m) {
  S->m = m;
  for (int i = 0; i < m; ++i) S->ids[i] = ids[i];
  S->dependent = 0;
  S->well = 1;
  const ipt *p0 = &cl->P[ids[0]];
  bi t, u;
  if (m == 1) {
    bi_zero(&S->lnum);
    bi_set_i64(&S->lden, 1);
    for (int a = 0; a < 3; ++a) bi_copy(&S->cnum[a], &p0->c[a]);
    bi_set_i64(&S->cden, 1);
    return 1;
  }
  vsub(S->a, &cl->P[ids[1]], p0);
  if (m == 2) {
    vdot(&S->lnum, S->a, S->a);
    bi_set_i64(&S->lden, 4);
    for (int a = 0; a < 3; ++a) bi_add(&S->cnum[a], &p0->c[a], &cl->P[ids[1]].c[a]);
    bi_set_i64(&S->cden, 2);
    return 1;
  }
  vsub(S->b, &cl->P[ids[2]], p0);
  if (m == 3) {
    bi aa, bb, ab, amb[3], q;
    vdot(&aa, S->a, S->a);
    vdot(&bb, S->b, S->b);
    vdot(&ab, S->a, S->b);
    bi_mul(&t, &aa, &bb);
    bi_mul(&u, &ab, &ab);
    bi_sub(&S->D, &t, &u);
    if (S->D.s == 0) { S->dependent = 1; return 0; }
    bi_sub(&t, &aa, &ab);
    bi_mul(&S->al, &bb, &t);
    bi_sub(&t, &bb, &ab);
    bi_mul(&S->be, &aa, &t);
    bi_add(&t, &S->al, &S->be);
    bi_add(&u, &S->D, &S->D);
    bi_sub(&u, &u, &t);
    S->well = S->al.s > 0 && S->be.s > 0 && u.s > 0;
    for (int a = 0; a < 3; ++a) bi_sub(&amb[a], &S->a[a], &S->b[a]);
    vdot(&q, amb, amb);
    bi_mul(&t, &aa, &bb);
    bi_mul(&S->lnum, &t, &q);
    bi_mul_i64(&S->lden, &S->D, 4);
    /* centre = p0 + (al a + be b) / (2 D) */
    bi_add(&S->cden, &S->D, &S->D);
    for (int a = 0; a < 3; ++a) {
      bi_mul(&t, &S->al, &S->a[a]);
      bi_mul(&u, &S->be, &S->b[a]);
      bi_add(&t, &t, &u);
      bi_mul(&u, &p0->c[a], &S->cden);
      bi_add(&S->cnum[a], &t, &u);
    }
    return 1;
  }
  vsub(S->d, &cl->P[ids[3]], p0);
  {
    bi bxd[3], dxa[3], axb[3], aa, bb, dd, nv, nw, nz, nu, tmp[3];
    vcross(bxd, S->b, S->d);
    vcross(dxa, S->d, S->a);
    vcross(axb, S->a, S->b);
    vdot(&S->V, S->a, bxd);
    if (S->V.s == 0) { S->dependent = 1; return 0; }
    vdot(&aa, S->a, S->a);
    vdot(&bb, S->b, S->b);
    vdot(&dd, S->d, S->d);
    for (int a = 0; a < 3; ++a) {
      bi_mul(&tmp[0], &aa, &bxd[a]);
      bi_mul(&tmp[1], &bb, &dxa[a]);
      bi_mul(&tmp[2], &dd, &axb[a]);
      bi_add(&t, &tmp[0], &tmp[1]);
      bi_add(&S->N[a], &t, &tmp[2]);
    }
    vdot(&nv, S->N, bxd);
    vdot(&nw, S->N, dxa);
    vdot(&nz, S->N, axb);
    bi_mul(&t, &S->V, &S->V);
    bi_add(&nu, &t, &t);
    bi_sub(&nu, &nu, &nv);
    bi_sub(&nu, &nu, &nw);
    bi_sub(&nu, &nu, &nz);
    S->well = nv.s > 0 && nw.s > 0 && nz.s > 0 && nu.s > 0;
    vdot(&S->lnum, S->N, S->N);
    bi_mul(&t, &S->V, &S->V);
    bi_mul_i64(&S->lden, &t, 4);
    /* centre = p0 + N / (2V), normalised to a positive denominator */
    bi_add(&S->cden, &S->V, &S->V);
    for (int a = 0; a < 3; ++a) {
      bi_mul(&u, &p0->c[a], &S->cden);
      bi_add(&S->cnum[a], &S->N[a], &u);
    }
    if (S->cden.s < 0) {
      bi_neg(&S->cden, &S->cden);
      for (int a = 0; a < 3; ++a) bi_neg(&S->cnum[a], &S->cnum[a]);
    }
  }
  return 1;
}

/* sign of the power of x with respect to the support's circumsphere:
 * -1 strictly inside, 0 on the sphere, +1 outside */
static int support_side(const sup_t *S, const ecloud *cl, int64_t x) {
  const ipt *p0 = &cl->P[S->ids[0]];
  const ipt *px = &cl->P[x];
  bi p[3], t, u;
  if (S->m == 1) {
    for (int a = 0; a < 3; ++a)
      if (bi_cmp(&p0->c[a], &px->c[a])) return 1;
    return 0;
  }
  vsub(p, px, p0);
  if (S->m == 2) {
    bi q[3];
    vsub(q, px, &cl->P[S->ids[1]]);
    vdot(&t, p, q);
    return t.s;
  }
  if (S->m == 3) {
    bi pp, pa, pb;
    vdot(&pp, p, p);
    vdot(&pa, p, S->a);
    vdot(&pb, p, S->b);
    bi_mul(&t, &S->D, &pp);
    bi_mul(&u, &S->al, &pa);
    bi_sub(&t, &t, &u);
    bi_mul(&u, &S->be, &pb);
    bi_sub(&t, &t, &u);
    return t.s; /* D > 0 */
  }
  {
    bi pp, pn;
    vdot(&pp, p, p);
    vdot(&pn, p, S->N);
    bi_mul(&t, &S->V, &pp);
    bi_sub(&t, &t, &pn);
    return t.s * S->V.s;
  }
}

/* ------------------------------------------------------------------------ */
/* exhaustive critical catalogue                                             */
/* reference: reference/morsehgp3d_oracle/catalog.py build_critical_catalog  */
/* and morsehgp3d/src/cpu/hierarchy/critical_catalog.cpp (exhaustive        */
/* supports 1..4, well-centred, shell == support, closed rank <= s_max,      */
/* relevant extra-shell equalities kept as degeneracies)                     */
/* ------------------------------------------------------------------------ */

#define MAX_INTERIOR 16

typedef struct {
  int m;
  int64_t sup[4];
  int rank;
  int nint;
  int64_t interior[MAX_INTERIOR];
  int degenerate;      /* 1: extra shell (not an event) */
  int shell_count;
  int64_t shell_witness; /* least extra-shell id */
  bi lnum, lden;
  double level;        /* rounded value in input units (for sorting only) */
} cat_rec;

static cat_rec *g_cat;
static int64_t g_cat_n, g_cat_cap;
static ecloud g_cat_cl;
static int g_cat_valid;

static void cat_push(const cat_rec *r) {
  if (g_cat_n == g_cat_cap) {
    g_cat_cap = g_cat_cap ? 2 * g_cat_cap : 1024;
    g_cat = realloc(g_cat, sizeof(cat_rec) * (size_t)g_cat_cap);
  }
  g_cat[g_cat_n++] = *r;
}

static void next_comb(int64_t *c, int m, int64_t n, int *done) {
  int i = m - 1;
  while (i >= 0 && c[i] == n - m + i) i--;
  if (i < 0) { *done = 1; return; }
  c[i]++;
  for (int j = i + 1; j < m; ++j) c[j] = c[j - 1] + 1;
}

/* Builds the catalogue for s_max = min(min(k_max, n) + 1, n).  Returns the
 * number of records (events and relevant degeneracies) or a negative error. */
int64_t or_catalog_build(const double *xyz, int64_t n, int k_max, int max_support) {
  if (g_cat_valid) { ecloud_free(&g_cat_cl); g_cat_valid = 0; }
  g_cat_n = 0;
  if (n <= 0 || k_max < 1) return OR_ARGUMENT;
  int rc = ecloud_init(&g_cat_cl, xyz, n);
  if (rc) return rc;
  g_cat_valid = 1;
  int64_t keff = k_max < n ? k_max : n;
  int64_t smax = keff + 1 < n ? keff + 1 : n;
  sup_t *S = malloc(sizeof(sup_t));
  cat_rec r;
  int topm = max_support < 4 ? max_support : 4;
  for (int m = 1; m <= topm && m <= n; ++m) {
    if (m > smax) break; /* closed rank >= |U| > s_max: above rank, never relevant */
    int64_t c[4];
    for (int i = 0; i < m; ++i) c[i] = i;
    int done = 0;
    while (!done) {
      if (support_analyze(S, &g_cat_cl, c, m) && S->well) {
        int nint = 0, nshell = 0;
        int64_t wit = -1;
        int over = 0;
        for (int64_t x = 0; x < n && !over; ++x) {
          int onsup = 0;
          for (int i = 0; i < m; ++i) onsup |= (c[i] == x);
          if (onsup) continue;
          int s = support_side(S, &g_cat_cl, x);
          if (s < 0) {
            if (nint < MAX_INTERIOR) r.interior[nint] = x;
            nint++;
            if (nint + m > smax) over = 1;
          } else if (s == 0) {
            nshell++;
            if (wit < 0) wit = x;
          }
        }
        if (!over) {
          r.m = m;
          for (int i = 0; i < 4; ++i) r.sup[i] = i < m ? c[i] : -1;
          r.nint = nint;
          r.rank = nint + m;
          r.degenerate = nshell > 0;
          r.shell_count = m + nshell;
          r.shell_witness = wit;
          bi_copy(&r.lnum, &S->lnum);
          bi_copy(&r.lden, &S->lden);
          r.level = bi_to_double(&S->lnum) / bi_to_double(&S->lden) * ldexp(1.0, 2 * g_cat_cl.e0);
          cat_push(&r);
        }
      }
      next_comb(c, m, n, &done);
    }
  }
  free(S);
  return g_cat_n;
}

/* record accessor; level returned as reduced decimal num/den in input units */
int or_catalog_get(int64_t i, int *m, int64_t *sup, int *rank, int *nint, int64_t *interior,
                   int *degenerate, int *shell_count, int64_t *shell_witness, double *level,
                   char *num, char *den, int cap) {
  if (i < 0 || i >= g_cat_n) return OR_ARGUMENT;
  cat_rec *r = &g_cat[i];
  *m = r->m;
  for (int k = 0; k < 4; ++k) sup[k] = r->sup[k];
  *rank = r->rank;
  *nint = r->nint;
  for (int k = 0; k < r->nint && k < MAX_INTERIOR; ++k) interior[k] = r->interior[k];
  *degenerate = r->degenerate;
  *shell_count = r->shell_count;
  *shell_witness = r->shell_witness;
  *level = r->level;
  if (num && den) {
    bi nn, dd, g;
    bi_copy(&nn, &r->lnum);
    bi_copy(&dd, &r->lden);
    int sh = 2 * g_cat_cl.e0;
    if (sh > 0) bi_shl(&nn, &nn, sh); else if (sh < 0) bi_shl(&dd, &dd, -sh);
    if (nn.s == 0) { bi_set_i64(&dd, 1); }
    else {
      bi_gcd(&g, &nn, &dd);
      bi_divmod(&nn, NULL, &nn, &g);
      bi_divmod(&dd, NULL, &dd, &g);
    }
    if (bi_to_dec(&nn, num, cap) < 0 || bi_to_dec(&dd, den, cap) < 0) return OR_CAPACITY;
  }
  return OR_OK;
}

/* exact comparison of the squared levels of two records */
int or_catalog_level_cmp(int64_t i, int64_t j) {
  bi x, y;
  bi_mul(&x, &g_cat[i].lnum, &g_cat[j].lden);
  bi_mul(&y, &g_cat[j].lnum, &g_cat[i].lden);
  return bi_cmp(&x, &y);
}

/* ------------------------------------------------------------------------ */
/* exact helpers exposed for the small-n tower oracle                        */
/* ------------------------------------------------------------------------ */

/* side of point x against the circumsphere (centre in the affine hull) of the
 * given support, on the catalogue's cloud.  Returns -1/0/1, or 2 if dependent. */
int or_support_side(const int64_t *ids, int m, int64_t x) {
  sup_t *S = malloc(sizeof(sup_t));
  int r = support_analyze(S, &g_cat_cl, ids, m) ? support_side(S, &g_cat_cl, x) : 2;
  free(S);
  return r;
}

/* well-centredness and level of a support; level as reduced num/den strings */
int or_support_info(const int64_t *ids, int m, int *well, char *num, char *den, int cap) {
  sup_t *S = malloc(sizeof(sup_t));
  if (!support_analyze(S, &g_cat_cl, ids, m)) { free(S); return 2; }
  *well = S->well;
  bi nn, dd, g;
  bi_copy(&nn, &S->lnum);
  bi_copy(&dd, &S->lden);
  int sh = 2 * g_cat_cl.e0;
  if (sh > 0) bi_shl(&nn, &nn, sh); else if (sh < 0) bi_shl(&dd, &dd, -sh);
  if (nn.s == 0) bi_set_i64(&dd, 1);
  else { bi_gcd(&g, &nn, &dd); bi_divmod(&nn, NULL, &nn, &g); bi_divmod(&dd, NULL, &dd, &g); }
  bi_to_dec(&nn, num, cap);
  bi_to_dec(&dd, den, cap);
  free(S);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* anchored exhaustive catalogue (CPU baseline of the timing harness)        */
/* Enumerates, exactly, every support whose least PointId is the anchor among */
/* the points within `radius` of it, and classifies the neighbourhood.  A    */
/* FP64 prefilter skips supports whose circumradius exceeds radius / 2 or    */
/* that are clearly not well-centred; every kept decision is exact.          */
/* ------------------------------------------------------------------------ */

static ecloud g_anc_cl;
static int g_anc_valid;

int or_anchor_prepare(const double *xyz, int64_t n) {
  if (g_anc_valid) { ecloud_free(&g_anc_cl); g_anc_valid = 0; }
  int rc = ecloud_init(&g_anc_cl, xyz, n);
  if (rc == OR_OK) g_anc_valid = 1;
  return rc;
}

static int f_circum(const double *X, const int64_t *ids, int m, double *r2, int *well) {
  const double *p = X + 3 * ids[0];
  double a[3], b[3], d[3];
  for (int k = 0; k < 3; ++k) a[k] = X[3 * ids[1] + k] - p[k];
  if (m == 2) { *r2 = 0.25 * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); *well = 1; return 1; }
  for (int k = 0; k < 3; ++k) b[k] = X[3 * ids[2] + k] - p[k];
  double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2], bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  if (m == 3) {
    double D = aa * bb - ab * ab;
    if (!(D > 1e-12 * aa * bb)) return 0;
    double al = bb * (aa - ab), be = aa * (bb - ab);
    double c[3];
    for (int k = 0; k < 3; ++k) c[k] = (al * a[k] + be * b[k]) / (2 * D);
    *r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    *well = al > -1e-9 * fabs(al) && be > -1e-9 * fabs(be) && 2 * D - al - be > -1e-9 * (2 * D);
    return 1;
  }
  for (int k = 0; k < 3; ++k) d[k] = X[3 * ids[3] + k] - p[k];
  double bxd[3] = {b[1] * d[2] - b[2] * d[1], b[2] * d[0] - b[0] * d[2], b[0] * d[1] - b[1] * d[0]};
  double dxa[3] = {d[1] * a[2] - d[2] * a[1], d[2] * a[0] - d[0] * a[2], d[0] * a[1] - d[1] * a[0]};
  double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  double V = a[0] * bxd[0] + a[1] * bxd[1] + a[2] * bxd[2];
  if (fabs(V) < 1e-12 * sqrt(aa * bb * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))) return 0;
  double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double N[3];
  for (int k = 0; k < 3; ++k) N[k] = aa * bxd[k] + bb * dxa[k] + dd * axb[k];
  *r2 = (N[0] * N[0] + N[1] * N[1] + N[2] * N[2]) / (4 * V * V);
  double nv = N[0] * bxd[0] + N[1] * bxd[1] + N[2] * bxd[2];
  double nw = N[0] * dxa[0] + N[1] * dxa[1] + N[2] * dxa[2];
  double nz = N[0] * axb[0] + N[1] * axb[1] + N[2] * axb[2];
  double tol = 1e-9 * (fabs(nv) + fabs(nw) + fabs(nz) + 2 * V * V);
  *well = nv > -tol && nw > -tol && nz > -tol && 2 * V * V - nv - nw - nz > -tol;
  return 1;
}

int64_t or_anchor_catalog(int k_max, int64_t anchor, double radius) {
  if (!g_anc_valid) return OR_ARGUMENT;
  const ecloud *cl = &g_anc_cl;
  const double *X = cl->X;
  const int64_t n = cl->n;
  int64_t keff = k_max < n ? k_max : n;
  int64_t smax = keff + 1 < n ? keff + 1 : n;
  int64_t *nb = malloc(sizeof(int64_t) * (size_t)n);
  int64_t m = 0;
  for (int64_t j = 0; j < n; ++j) {
    double s = 0;
    for (int k = 0; k < 3; ++k) { double t = X[3 * j + k] - X[3 * anchor + k]; s += t * t; }
    if (s <= radius * radius && j != anchor) nb[m++] = j;
  }
  sup_t *S = malloc(sizeof(sup_t));
  int64_t events = 0;
  const double rmax2 = 0.25 * radius * radius;
  int64_t ids[4];
  ids[0] = anchor;
  for (int sz = 2; sz <= 4 && sz <= smax; ++sz) {
    int64_t c[3] = {0, 1, 2};
    int k = sz - 1;
    if (m < k) break;
    int done = 0;
    while (!done) {
      int ok = 1;
      for (int q = 0; q < k; ++q) { ids[q + 1] = nb[c[q]]; ok &= nb[c[q]] > anchor; }
      double r2; int well;
      if (ok && f_circum(X, ids, sz, &r2, &well) && well && r2 <= rmax2 &&
          support_analyze(S, cl, ids, sz) && S->well) {
        int nint = 0, shell = 0;
        for (int64_t t = 0; t < m && nint + sz <= smax; ++t) {
          int64_t x = nb[t];
          int on = 0;
          for (int q = 1; q < sz; ++q) on |= ids[q] == x;
          if (on) continue;
          int sd = support_side(S, cl, x);
          if (sd < 0) nint++; else if (sd == 0) shell++;
        }
        if (nint + sz <= smax && shell == 0) events++;
      }
      next_comb(c, k, m, &done);
    }
  }
  free(S);
  free(nb);
  return events;
}

/* ------------------------------------------------------------------------ */
/* anchored exhaustive records (parity of the device catalogue at scale)     */
/* Every support U of 2..4 points whose least PointId is `anchor`, all of    */
/* whose points lie within `radius` of it, classified exactly against every  */
/* point within `radius` (a sphere through the anchor of radius r lies in    */
/* the ball of radius 2r around it, so the classification of supports with  */
/* r <= radius / 2 is complete).  Same record definition as                  */
/* or_catalog_build: well-centred, closed rank <= s_max, extra shell points  */
/* flag the record as degenerate.  FP64 filters only skip work they can      */
/* certify is not a record (well-conditioned cases, margins far above their  */
/* rounding); everything else is decided exactly.  Reentrant: several        */
/* threads may run it at once on the cloud of or_anchor_prepare.             */
/* ------------------------------------------------------------------------ */

#define ANC_INTERIOR 10 /* s_max - 2 for s_max <= 11 */

typedef struct {
  const ecloud *cl;
  int64_t anchor, smax;
  const int64_t *nb;  /* every point within radius except the anchor */
  int64_t nnb;
  const double *rel;  /* offsets of nb from the anchor, 3 per point */
  double rmax2;
  sup_t *S;
  int64_t count, cap;
  int32_t *m_out, *rank_out, *nint_out, *deg_out, *shell_out;
  int64_t *sup_out, *int_out;
  char *num_out, *den_out;
  int strcap;
  int err;
} anc_ctx;

/* FP64 circumsphere of the anchor (origin) and 1..3 offsets: centre offset and
 * squared radius; returns 1 when the support is well conditioned (the values
 * are then accurate to far better than 1e-9 relative), 0 otherwise. */
static int anc_fp_sphere(const double *a, const double *b, const double *d, int m, double c[3], double *r2,
                         int *clearly_not_well) {
  *clearly_not_well = 0;
  if (m == 2) {
    for (int k = 0; k < 3; ++k) c[k] = 0.5 * a[k];
    *r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    return 1;
  }
  double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2], bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  if (m == 3) {
    double D = aa * bb - ab * ab;
    if (!(D > 1e-6 * aa * bb)) return 0;
    double al = bb * (aa - ab), be = aa * (bb - ab), ga = 2 * D - al - be;
    for (int k = 0; k < 3; ++k) c[k] = (al * a[k] + be * b[k]) / (2 * D);
    *r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    double tol = 1e-9 * aa * bb;
    *clearly_not_well = al < -tol || be < -tol || ga < -tol;
    return 1;
  }
  double bxd[3] = {b[1] * d[2] - b[2] * d[1], b[2] * d[0] - b[0] * d[2], b[0] * d[1] - b[1] * d[0]};
  double dxa[3] = {d[1] * a[2] - d[2] * a[1], d[2] * a[0] - d[0] * a[2], d[0] * a[1] - d[1] * a[0]};
  double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  double V = a[0] * bxd[0] + a[1] * bxd[1] + a[2] * bxd[2];
  double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double scale = sqrt(aa * bb * dd);
  if (!(fabs(V) > 1e-6 * scale)) return 0;
  double N[3];
  for (int k = 0; k < 3; ++k) N[k] = aa * bxd[k] + bb * dxa[k] + dd * axb[k];
  for (int k = 0; k < 3; ++k) c[k] = N[k] / (2 * V);
  *r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  /* barycentric coordinates of the centre times 2 V^2 */
  double nv = N[0] * bxd[0] + N[1] * bxd[1] + N[2] * bxd[2];
  double nw = N[0] * dxa[0] + N[1] * dxa[1] + N[2] * dxa[2];
  double nz = N[0] * axb[0] + N[1] * axb[1] + N[2] * axb[2];
  double n0 = 2 * V * V - nv - nw - nz;
  double tol = 1e-9 * scale * scale;
  *clearly_not_well = nv < -tol || nw < -tol || nz < -tol || n0 < -tol;
  return 1;
}

/* FP64 count of the points certainly strictly inside a well-conditioned sphere;
 * stops past `limit`. */
static int64_t anc_fp_inside(const anc_ctx *A, const double c[3], double r2, int64_t limit) {
  int64_t cnt = 0;
  const double lim2 = r2 * (1.0 - 1e-7);
  for (int64_t t = 0; t < A->nnb; ++t) {
    const double *p = A->rel + 3 * t;
    double x = p[0] - c[0], y = p[1] - c[1], z = p[2] - c[2];
    if (x * x + y * y + z * z < lim2 && ++cnt > limit) break;
  }
  return cnt;
}

static void anc_exact(anc_ctx *A, const int64_t *ids, int m) {
  if (A->err) return;
  sup_t *S = A->S;
  if (!support_analyze(S, A->cl, ids, m) || !S->well) return;
  int64_t nint = 0, nshell = 0, inter[ANC_INTERIOR];
  for (int64_t t = 0; t < A->nnb; ++t) {
    int64_t x = A->nb[t];
    int on = 0;
    for (int q = 1; q < m; ++q) on |= ids[q] == x;
    if (on) continue;
    int sd = support_side(S, A->cl, x);
    if (sd < 0) {
      if (nint < ANC_INTERIOR) inter[nint] = x;
      if (++nint + m > A->smax) return;
    } else if (sd == 0) {
      nshell++;
    }
  }
  int64_t i = A->count++;
  if (i >= A->cap) return;
  A->m_out[i] = m;
  for (int q = 0; q < 4; ++q) A->sup_out[4 * i + q] = q < m ? ids[q] : -1;
  A->rank_out[i] = (int32_t)(nint + m);
  A->nint_out[i] = (int32_t)nint;
  /* interior ascending */
  for (int64_t a = 1; a < nint; ++a)
    for (int64_t b = a; b > 0 && inter[b - 1] > inter[b]; --b) { int64_t t = inter[b]; inter[b] = inter[b - 1]; inter[b - 1] = t; }
  for (int q = 0; q < ANC_INTERIOR; ++q) A->int_out[ANC_INTERIOR * i + q] = q < nint ? inter[q] : -1;
  A->deg_out[i] = nshell > 0;
  A->shell_out[i] = (int32_t)(m + nshell);
  bi nn, dd, g;
  bi_copy(&nn, &S->lnum);
  bi_copy(&dd, &S->lden);
  int sh = 2 * A->cl->e0;
  if (sh > 0) bi_shl(&nn, &nn, sh); else if (sh < 0) bi_shl(&dd, &dd, -sh);
  if (nn.s == 0) bi_set_i64(&dd, 1);
  else { bi_gcd(&g, &nn, &dd); bi_divmod(&nn, NULL, &nn, &g); bi_divmod(&dd, NULL, &dd, &g); }
  if (bi_to_dec(&nn, A->num_out + (size_t)i * A->strcap, A->strcap) < 0 ||
      bi_to_dec(&dd, A->den_out + (size_t)i * A->strcap, A->strcap) < 0)
    A->err = OR_CAPACITY;
}

/* Returns the number of records (written up to cap; call again with a larger
 * cap when it exceeds it) or a negative error.  Per record: support size,
 * 4 ids (-1 padded), rank, interior count, 10 interior ids (-1 padded),
 * degenerate flag, shell count, and the squared radius as reduced decimal
 * num / den strings of strcap bytes each. */
int64_t or_anchor_records(int k_max, int64_t anchor, double radius, int64_t cap, int32_t *m_out,
                          int64_t *sup_out, int32_t *rank_out, int32_t *nint_out, int64_t *int_out,
                          int32_t *deg_out, int32_t *shell_out, char *num_out, char *den_out, int strcap) {
  if (!g_anc_valid) return OR_ARGUMENT;
  const ecloud *cl = &g_anc_cl;
  const double *X = cl->X;
  const int64_t n = cl->n;
  if (anchor < 0 || anchor >= n || !(radius > 0)) return OR_ARGUMENT;
  anc_ctx A;
  memset(&A, 0, sizeof(A));
  A.cl = cl;
  A.anchor = anchor;
  int64_t keff = k_max < n ? k_max : n;
  A.smax = keff + 1 < n ? keff + 1 : n;
  if (A.smax > 11) return OR_ARGUMENT;
  int64_t *nb = malloc(sizeof(int64_t) * (size_t)n);
  double *rel = malloc(sizeof(double) * 3 * (size_t)n);
  int64_t *cand = malloc(sizeof(int64_t) * (size_t)n);
  int64_t *cidx = malloc(sizeof(int64_t) * (size_t)n);
  A.S = malloc(sizeof(sup_t));
  if (!nb || !rel || !cand || !cidx || !A.S) { free(nb); free(rel); free(cand); free(cidx); free(A.S); return OR_CAPACITY; }
  const double R2 = radius * radius * (1.0 + 1e-9);
  int64_t m = 0, mc = 0;
  for (int64_t j = 0; j < n; ++j) {
    if (j == anchor) continue;
    double d[3], s = 0;
    for (int k = 0; k < 3; ++k) { d[k] = X[3 * j + k] - X[3 * anchor + k]; s += d[k] * d[k]; }
    if (s > R2) continue;
    nb[m] = j;
    for (int k = 0; k < 3; ++k) rel[3 * m + k] = d[k];
    if (j > anchor) { cand[mc] = j; cidx[mc] = m; mc++; }
    m++;
  }
  A.nb = nb; A.nnb = m; A.rel = rel;
  A.rmax2 = 0.25 * radius * radius * (1.0 + 1e-6);
  A.cap = cap;
  A.m_out = m_out; A.sup_out = sup_out; A.rank_out = rank_out; A.nint_out = nint_out; A.int_out = int_out;
  A.deg_out = deg_out; A.shell_out = shell_out; A.num_out = num_out; A.den_out = den_out; A.strcap = strcap;
  const double pair2 = radius * radius * (1.0 + 1e-6);
  int64_t ids[4];
  ids[0] = anchor;
  for (int64_t i = 0; i < mc && A.smax >= 2; ++i) {
    const double *a = rel + 3 * cidx[i];
    double c[3], r2;
    int notwell;
    anc_fp_sphere(a, NULL, NULL, 2, c, &r2, &notwell);
    if (r2 > A.rmax2) continue;
    if (anc_fp_inside(&A, c, r2, A.smax - 2) > A.smax - 2) continue;
    ids[1] = cand[i];
    anc_exact(&A, ids, 2);
  }
  for (int64_t i = 0; i < mc && A.smax >= 3; ++i) {
    const double *a = rel + 3 * cidx[i];
    for (int64_t j = i + 1; j < mc; ++j) {
      const double *b = rel + 3 * cidx[j];
      double ab2 = 0;
      for (int k = 0; k < 3; ++k) ab2 += (a[k] - b[k]) * (a[k] - b[k]);
      if (ab2 > pair2) continue;  /* two points of a sphere of radius <= radius/2 */
      double c[3], r2;
      int notwell;
      int good = anc_fp_sphere(a, b, NULL, 3, c, &r2, &notwell);
      /* a tetrahedron's circumradius is at least that of each face */
      if (good && r2 > A.rmax2) continue;
      ids[1] = cand[i];
      ids[2] = cand[j];
      if (!(good && (notwell || anc_fp_inside(&A, c, r2, A.smax - 3) > A.smax - 3))) anc_exact(&A, ids, 3);
      if (A.smax < 4) continue;
      for (int64_t l = j + 1; l < mc; ++l) {
        const double *d = rel + 3 * cidx[l];
        double ad2 = 0, bd2 = 0;
        for (int k = 0; k < 3; ++k) { ad2 += (a[k] - d[k]) * (a[k] - d[k]); bd2 += (b[k] - d[k]) * (b[k] - d[k]); }
        if (ad2 > pair2 || bd2 > pair2) continue;
        double c4[3], r24;
        int notwell4;
        int good4 = anc_fp_sphere(a, b, d, 4, c4, &r24, &notwell4);
        if (good4 && (r24 > A.rmax2 || notwell4)) continue;
        if (good4 && anc_fp_inside(&A, c4, r24, A.smax - 4) > A.smax - 4) continue;
        ids[3] = cand[l];
        anc_exact(&A, ids, 4);
      }
    }
  }
  free(nb); free(rel); free(cand); free(cidx); free(A.S);
  if (A.err) return A.err;
  return A.count;
}

/* ------------------------------------------------------------------------ */
/* exact minimum spanning forest of a given edge set under kappa=(d^2,u,v)   */
/* (Kruskal as in or_emst, over candidate edges instead of every pair; the   */
/* EMST when the candidates contain it, e.g. a k-nearest-neighbour graph)    */
/* ------------------------------------------------------------------------ */

/* cand: 2m ids (any order inside a pair, duplicates allowed); forest edges
 * (u < v) are written in kappa order; returns their number or an error. */
int64_t or_emst_graph(const double *xyz, int64_t n, const int64_t *cand, int64_t m, int64_t *edges) {
  if (n < 1 || m < 0) return OR_ARGUMENT;
  ecloud cl;
  int rc = ecloud_init(&cl, xyz, n);
  if (rc) return rc;
  fedge *E = malloc(sizeof(fedge) * (size_t)(m ? m : 1));
  int64_t *par = malloc(sizeof(int64_t) * (size_t)n);
  if (!E || !par) { free(E); free(par); ecloud_free(&cl); return OR_CAPACITY; }
  int64_t t = 0;
  for (int64_t i = 0; i < m; ++i) {
    int64_t u = cand[2 * i], v = cand[2 * i + 1];
    if (u == v || u < 0 || v < 0 || u >= n || v >= n) continue;
    if (u > v) { int64_t s = u; u = v; v = s; }
    E[t].f = fdist2(xyz, u, v); E[t].u = u; E[t].v = v; t++;
  }
  g_emst_cl = &cl;
  qsort(E, (size_t)t, sizeof(fedge), edge_cmp);
  for (int64_t i = 0; i < n; ++i) par[i] = i;
  int64_t k = 0;
  for (int64_t i = 0; i < t && k < n - 1; ++i) {
    if (i > 0 && E[i].u == E[i - 1].u && E[i].v == E[i - 1].v) continue;
    int64_t a = uf_find(par, E[i].u), b = uf_find(par, E[i].v);
    if (a == b) continue;
    par[a] = b;
    edges[2 * k] = E[i].u;
    edges[2 * k + 1] = E[i].v;
    k++;
  }
  free(par);
  free(E);
  ecloud_free(&cl);
  return k;
}
